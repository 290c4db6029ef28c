"""GPU runs on the reference's own G1 geometry and of the calcium_sparks
program end to end (output file checked against the oracle)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import csparks as cs
import dyad
from helpers import load_geometry_fixture, load_model_fixture, species_close, with_serca
from oracle_py import Oracle

pytestmark = pytest.mark.gpu
CLI = os.path.join(cs.BUILD, "calcium_sparks")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("h,steps,fuse", [(12.0, 200, 0), (12.0, 200, 1), (6.0, 30, 0), (6.0, 31, 1)])
def test_g1_reference_geometry_fp64(h, steps, fuse):
    """The shipped G1 geometry (84^3 at 12 nm, 31416 SERCA pairs, 50 RyRs)
    with the shipped sr1100/ryr55 species file, RyR 24 forced open and
    closed at 0.5 ms: GPU vs oracle to 1e-10 per voxel, in one-step and in
    fused passes (more flux rows than one bridge block takes, so every
    bridge phase runs as a launch of many blocks)."""
    g = load_geometry_fixture("g1_nonP_0")
    m = load_model_fixture("params_sr1100_ryr55")
    opts = cs.Options(h=h, seed=11, open={"ryr": [24]}, t_close={"ryr": 0.5})
    _say(f"case: G1 h={h} steps={steps} fuse={fuse}")
    orc = Oracle(g, m, opts)
    sim = cs.Simulation(g, m, opts)
    if fuse:
        ok, why = sim.fused_available()
        assert ok, why
    sim.set_launch_config(fuse=fuse)
    sim.step(steps)
    if fuse:
        _, _, passes, fused = sim.time_passes(2)
        orc.step(2)
        assert fused == passes == 1
    orc.step(steps)
    for s in range(orc.ns):
        ok, err = species_close(sim.species(s), orc.species(s), 1e-10)
        assert ok, f"species {s}: {err:.3e}"
    a, b = sim.stats(), orc.stats()
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-10, atol=0)
    sim.close()
    orc.close()


def _inputs(tmp_path, d):
    gp, mp = str(tmp_path / "geom.h5"), str(tmp_path / "species.h5")
    cs.write_geometry(gp, d.geometry)
    cs.write_species(mp, d.model)
    return gp, mp


def _run_cli(args, cwd):
    _say("cli: " + " ".join(args))
    r = subprocess.run([CLI] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def test_cli_end_to_end_matches_oracle(tmp_path):
    d = with_serca(dyad.small_dyad(seed=17))
    gp, mp = _inputs(tmp_path, d)
    r = _run_cli(["-g", gp, "-m", mp, "-t", "0.01", "-o", "0.002", "-S", "3", "-C", "ryr", "0.005",
                  "-c", "run", "-s", "Ca", "Fluo", "-z", "30", "--all_data", "Ca", "CMDN-CSQN",
                  "-O", "ryr", str(d.open_index), "-l", "x", "Fluo", "1", "-2"], tmp_path)
    out = str(tmp_path / "run.h5")
    opts = cs.Options(seed=3, t_close={"ryr": 0.005}, open={"ryr": [d.open_index]})
    orc = Oracle(d.geometry, d.model, opts)
    dt = orc.dt
    save = cs.host_lib().cs_save_interval(0.002, dt)
    nsteps = cs.host_lib().cs_count_steps(0.01, dt)
    last_out = nsteps // save
    # root attributes (species.c:2273-2281)
    assert list(cs.read_output_attr(out, "/", "N")) == list(orc.n)
    assert cs.read_output_attr(out, "/", "h")[0] == orc.info.h
    assert cs.read_output_attr_str(out, "/", "geometry_file") == gp
    ti = 0
    for k in range(last_out + 1):
        grp = f"/data_{k:07d}"
        if k > 0:
            orc.step(save)
            ti += save
        assert cs.read_output_attr(out, grp, "time_index")[0] == ti
        mn, mx, tot = orc.stats()
        for di, dn in enumerate(d.model.domain_names):
            for si, sn in enumerate(d.model.species_names):
                got = cs.read_output_attr(out, f"{grp}/{dn}/{sn}", "max")[0]
                assert got == pytest.approx(mx[di, si], rel=1e-10)
        np.testing.assert_array_equal(cs.read_output_dataset(out, f"{grp}/discrete_ryr"), orc.states(0))
    ca = cs.read_output_dataset(out, f"/data_{last_out:07d}/Ca_all_data")
    ok, err = species_close(ca, orc.species(0), 1e-10)
    assert ok, err
    sheet = cs.read_output_dataset(out, f"/data_{last_out:07d}/Fluo_z_sheet_5")
    np.testing.assert_allclose(sheet, orc.species(1)[:, :, 5], rtol=1e-10)
    # -l x Fluo 1 -2: offsets 1 on y and -2 on z (options.c:154-208)
    line = cs.read_output_dataset(out, f"/data_{last_out:07d}/linescan_Fluo_y_1_z_-2")
    want = orc.linescan(1, 0, (0, 1, -2), d.model.conv)
    np.testing.assert_allclose(line, want, rtol=1e-10, atol=1e-10 * float(np.max(np.abs(want))))
    assert f"{last_out}: time:" in r.stdout and "conservation" in r.stdout


def test_cli_mt_stream_matches_oracle(tmp_path):
    """--rng mt: the reference's Mersenne Twister draw order on the GPU."""
    import dataclasses
    d = dyad.small_dyad(seed=8)
    fast = dict(Kd_open=0.01, k_max_open=200.0, k_min_close=150.0, k_max_close=150.0)
    model = dataclasses.replace(d.model, ryr=dataclasses.replace(d.model.ryr, **fast),
                                p_ryr=dataclasses.replace(d.model.p_ryr, **fast))
    d = dyad.Dyad(d.geometry, model, d.open_index, d.description)
    gp, mp = _inputs(tmp_path, d)
    _run_cli(["-g", gp, "-m", mp, "-q", "-t", "0.005", "-o", "0.001", "-S", "21", "--rng", "mt",
              "-c", "mt"], tmp_path)
    out = str(tmp_path / "mt.h5")
    orc = Oracle(d.geometry, model, cs.Options(seed=21, rng="mt"))
    save = cs.host_lib().cs_save_interval(0.001, orc.dt)
    nsteps = cs.host_lib().cs_count_steps(0.005, orc.dt)
    changes = 0
    prev = orc.states(0).copy()
    for k in range(1, nsteps // save + 1):
        orc.step(save)
        np.testing.assert_array_equal(cs.read_output_dataset(out, f"/data_{k:07d}/discrete_ryr"), orc.states(0))
        np.testing.assert_array_equal(cs.read_output_dataset(out, f"/data_{k:07d}/discrete_p_ryr"), orc.states(1))
        changes += int(np.sum(orc.states(0) != prev))
        prev = orc.states(0).copy()
    assert changes > 0


def test_cli_abort_after_channels_close(tmp_path):
    """-X: stop tabort ms after an output first shows every channel closed
    (species.c:3384-3488).  At 6 nm dt = 0.95*36/(6*220000) ms, so -o 0.002
    saves every floor(0.002/dt) = 77 steps and -t 0.02 runs 772 steps: 11
    outputs with the one at t = 0.  -C 0.001 closes both channels at the
    gating step with ti dt + dt >= 0.001, ti = 38; the output at step 77 is
    the first to see them closed and starts the clock at 77 dt; the check
    after step k fires once (k - 77) dt >= 0.003, at k = 193, after the
    outputs at steps 0, 77 and 154 only."""
    d = dyad.small_dyad(seed=4)
    gp, mp = _inputs(tmp_path, d)
    common = ["-g", gp, "-m", mp, "-t", "0.02", "-o", "0.002", "-S", "1", "-O", "ryr", "0", "p_ryr", "0",
              "-C", "ryr", "0.001", "p_ryr", "0.001"]
    full = _run_cli(common + ["-c", "full"], tmp_path)
    ab = _run_cli(common + ["-c", "ab", "-X", "0.003"], tmp_path)
    n_full = full.stdout.count(": time:")
    n_ab = ab.stdout.count(": time:")
    assert (n_full, n_ab) == (11, 3), (n_full, n_ab)
