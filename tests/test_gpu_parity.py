"""GPU parity: the HIP path against the CPU oracle on the same inputs.

Deterministic FP64 runs must agree to 1e-10 relative per voxel (the bar of
a deterministic -O/-C run); stochastic runs draw the same Philox (or
Mersenne Twister) uniforms on both sides, so channel trajectories must be
identical and fields agree like deterministic runs; FP32 fields are held
to a stated per-species bound against the FP64 oracle.
"""
import sys

import numpy as np
import pytest

import csparks as cs
import dyad
from helpers import DEV_F32, deviation_close, rest_field, species_close, with_serca
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

RTOL_F64 = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _compare_fields(sim, orc, rtol):
    worst = 0.0
    for s in range(orc.ns):
        ok, err = species_close(sim.species(s), orc.species(s), rtol)
        worst = max(worst, err)
        assert ok, f"species {s}: max rel err {err:.3e} > {rtol}"
    return worst


def _run_pair(d, opts, nsteps, chunks=1, fuse=None):
    """GPU run and oracle over the same steps; fuse 0 / 1 selects one-step
    or fused two-step passes (None: the library's default)."""
    _say(f"case: {d.description} opts={opts} steps={nsteps} fuse={fuse}")
    orc = Oracle(d.geometry, d.model, opts)
    sim = cs.Simulation(d.geometry, d.model, opts)
    if fuse is not None:
        sim.set_launch_config(fuse=fuse)
    per = nsteps // chunks
    for _ in range(chunks):
        sim.step(per)
        orc.step(per)
    sim.synchronize()
    return sim, orc


def test_device_visible():
    _say(cs.device_info())


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_fp64_deterministic_small(fuse):
    d = dyad.small_dyad()
    opts = cs.Options(seed=7, open={"ryr": [d.open_index, 4]}, t_close={"ryr": 1.0}, tstop=1.0)
    sim, orc = _run_pair(d, opts, 300, chunks=3, fuse=fuse)
    worst = _compare_fields(sim, orc, RTOL_F64)
    _say(f"fp64 deterministic max rel err {worst:.2e}")
    np.testing.assert_array_equal(sim.channel_states(0), orc.states(0))
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_fp64_forced_close_and_substeps(fuse):
    """-C closes the forced-open channel mid-run; -T/-D substeps > 1."""
    d = dyad.small_dyad(seed=5)
    opts = cs.Options(seed=3, open={"ryr": [d.open_index]}, t_close={"ryr": 0.002, "p_ryr": 0.001},
                      dt_update_react=3.5e-5 * 3, dt_update_stoch=3.5e-5 * 2)
    sim, orc = _run_pair(d, opts, 240, chunks=2, fuse=fuse)
    assert sim.problem.reaction_substep > 1 and sim.problem.stochastic_substep > 1
    _compare_fields(sim, orc, RTOL_F64)
    assert sim.channel_states(0).sum() == 0
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_fp64_refined_geometry(fuse):
    """12 nm geometry run at -h 6: two subdivisions, discrete channels on
    the centre sub-voxel pair (species.c:1003-1020)."""
    d = dyad.small_dyad(n=(10, 9, 7), h=12.0, seed=11)
    opts = cs.Options(h=6.0, seed=1, open={"ryr": [0, 1]}, t_close={"ryr": 1.0})
    sim, orc = _run_pair(d, opts, 150, fuse=fuse)
    assert sim.n == (20, 18, 14)
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_fp64_serca_and_dirichlet(fuse):
    d = with_serca(dyad.small_dyad(seed=9))
    opts = cs.Options(seed=2, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0}, force_dirichlet=True)
    sim, orc = _run_pair(d, opts, 200, chunks=2, fuse=fuse)
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


def _with_model(d, **changes):
    import dataclasses
    return dyad.Dyad(d.geometry, dataclasses.replace(d.model, **changes), d.open_index, d.description)


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_reactions_in_slot_order_when_a_buffer_binds_another_species(fuse):
    """A cleft buffer whose free species is Fluo (not species 0) rules out
    the species-major form: every step takes the reference's slot-order
    path (reaction_SIMD, kernel_generic.cpp:397-509)."""
    d = dyad.small_dyad(seed=14)
    bufs = [list(b) for b in d.model.buffers]
    bufs[1][3] = cs.Buffer(4, 1, 0.0327, 0.0196, 70.0)
    d = _with_model(d, buffers=bufs)
    opts = cs.Options(seed=2, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    sim, orc = _run_pair(d, opts, 120, chunks=2, fuse=fuse)
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_reactions_overwrite_when_species0_does_not_diffuse(fuse):
    """Fluo diffuses fastest, so species 0 diffuses on a substep only; on
    the other steps every buffer slot overwrites species 0 with its own
    increment instead of accumulating (reaction_SIMD with update_species[0]
    unset, kernel_generic.cpp:413-427); the GPU must do the same."""
    d = dyad.small_dyad(seed=15)
    m = d.model
    sigma = [[{0: 22000.0, 1: 220000.0}.get(s, sg) for s, sg in zip(dl, sl)]
             for dl, sl in zip(m.diffusive, m.sigma)]
    d = _with_model(d, sigma=sigma)
    opts = cs.Options(seed=2, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    sim, orc = _run_pair(d, opts, 120, chunks=2, fuse=fuse)
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
@pytest.mark.parametrize("rng", ["philox", "mt"])
def test_stochastic_trajectories_match(rng, fuse):
    """Same uniforms on both sides => identical channel trajectories."""
    import dataclasses
    d = dyad.small_dyad(seed=4)
    # fast gating (rates ~100/ms) makes channels flicker within a few
    # hundred steps of the stable dt
    fast = dict(Kd_open=0.01, k_max_open=200.0, k_min_close=150.0, k_max_close=150.0)
    model = dataclasses.replace(d.model, ryr=dataclasses.replace(d.model.ryr, **fast),
                                p_ryr=dataclasses.replace(d.model.p_ryr, **fast))
    opts = cs.Options(seed=123, rng=rng, open={"ryr": [d.open_index]})
    orc = Oracle(d.geometry, model, opts)
    sim = cs.Simulation(d.geometry, model, opts)
    sim.set_launch_config(fuse=fuse)
    _say(f"case: stochastic {rng} fuse={fuse}")
    transitions = 0
    prev = orc.states(0).copy()
    for k in range(40):
        sim.step(5)
        orc.step(5)
        for f in range(2):
            a, b = sim.channel_states(f), orc.states(f)
            np.testing.assert_array_equal(a, b, err_msg=f"flux {f} after {5 * (k + 1)} steps")
        transitions += int(np.sum(orc.states(0) != prev))
        prev = orc.states(0).copy()
    assert transitions > 0, "no gating events happened; the test exercises nothing"
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


def test_fp32_against_fp64_oracle():
    d = with_serca(dyad.small_dyad(seed=6))
    opts64 = cs.Options(seed=8, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    opts32 = cs.Options(seed=8, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0}, precision="f32")
    orc = Oracle(d.geometry, d.model, opts64)
    sim = cs.Simulation(d.geometry, d.model, opts32)
    _say("case: fp32 vs fp64 oracle")
    sim.step(300)
    orc.step(300)
    errs = []
    for s in range(orc.ns):
        ok, rel, err, dev = deviation_close(sim.species(s), orc.species(s), rest_field(d, s), DEV_F32)
        errs.append(rel)
        assert ok, f"species {s}: fp32 error {err:.3e} > {DEV_F32} x max deviation {dev:.3e}"
    _say("fp32 error / max deviation from rest, per species: " + ", ".join(f"{e:.2e}" for e in errs))
    sim.close()


LAUNCHES = [dict(rows=8, zchunk=3, xcd_remap=1, prefetch=2, nontemporal=1, lds_tile=0, linear=0),
            dict(rows=1, zchunk=1, xcd_remap=0, prefetch=1, nontemporal=0, lds_tile=0, linear=0),
            dict(rows=2, zchunk=5, xcd_remap=1, prefetch=1, nontemporal=1, lds_tile=0, linear=0),
            dict(rows=4, zchunk=64, xcd_remap=0, prefetch=1, nontemporal=0, lds_tile=0, linear=0),
            dict(rows=4, zchunk=4, xcd_remap=1, prefetch=1, nontemporal=1, lds_tile=1, linear=0),
            dict(rows=8, zchunk=64, xcd_remap=0, prefetch=1, nontemporal=0, lds_tile=1, linear=0),
            dict(rows=1, zchunk=2, xcd_remap=0, prefetch=1, nontemporal=0, lds_tile=1, linear=0),
            dict(rows=2, zchunk=7, xcd_remap=1, prefetch=1, nontemporal=1, lds_tile=1, linear=0),
            dict(rows=4, zchunk=3, xcd_remap=1, prefetch=2, nontemporal=1, lds_tile=0, linear=1),
            dict(rows=1, zchunk=14, xcd_remap=0, prefetch=2, nontemporal=0, lds_tile=0, linear=1),
            dict(rows=8, zchunk=6, xcd_remap=0, prefetch=1, nontemporal=1, lds_tile=0, linear=1)]
LAUNCHES = [{**c, "yblock": 0} for c in LAUNCHES] + [
    dict(rows=1, zchunk=3, xcd_remap=1, prefetch=1, nontemporal=1, lds_tile=0, linear=0, yblock=2),
    dict(rows=2, zchunk=64, xcd_remap=0, prefetch=2, nontemporal=0, lds_tile=0, linear=0, yblock=4),
    dict(rows=1, zchunk=1, xcd_remap=0, prefetch=2, nontemporal=1, lds_tile=0, linear=0, yblock=4),
    dict(rows=2, zchunk=4, xcd_remap=3, prefetch=1, nontemporal=1, lds_tile=0, linear=1, yblock=0),
    dict(rows=1, zchunk=2, xcd_remap=16, prefetch=2, nontemporal=0, lds_tile=0, linear=0, yblock=0),
    dict(rows=1, zchunk=3, xcd_remap=5, prefetch=1, nontemporal=1, lds_tile=0, linear=0, yblock=2),
]


@pytest.mark.parametrize("cfg", LAUNCHES)
def test_launch_shapes_agree(cfg):
    """Every launch shape the autotuner may pick gives the same answer."""
    d = with_serca(dyad.small_dyad(seed=12))
    opts = cs.Options(seed=4, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    _say(f"case: launch {cfg}")
    orc = Oracle(d.geometry, d.model, opts)
    sim = cs.Simulation(d.geometry, d.model, opts)
    sim.set_launch_config(fuse=0, **cfg)   # one-step passes with this shape
    got = sim.launch_config()
    assert {k: got[k] for k in cfg} == {**cfg, "zchunk": min(cfg["zchunk"], sim.n[2])}
    sim.step(60)
    orc.step(60)
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


ROW_LAUNCHES = [dict(rows=1, zchunk=2, xcd_remap=0, prefetch=1, nontemporal=0, yblock=2),
                dict(rows=2, zchunk=4, xcd_remap=1, prefetch=1, nontemporal=1, yblock=4),
                dict(rows=3, zchunk=9, xcd_remap=0, prefetch=2, nontemporal=1, yblock=2),
                dict(rows=8, zchunk=5, xcd_remap=1, prefetch=2, nontemporal=0, yblock=4)]


@pytest.mark.parametrize("cfg", ROW_LAUNCHES)
def test_row_kernel_wide_grid(cfg):
    """The row-blocked kernel on a grid 150 wide (three wavefronts, the
    last one partial, x neighbours crossing wavefront and block edges) with
    13 rows (not a multiple of 2 or 4) and 9 planes."""
    d = with_serca(dyad.small_dyad(n=(150, 13, 9), seed=21, num_ryr=12, num_pryr=4))
    opts = cs.Options(seed=6, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    _say(f"case: row kernel {cfg} on {d.description}")
    orc = Oracle(d.geometry, d.model, opts)
    sim = cs.Simulation(d.geometry, d.model, opts)
    sim.set_launch_config(lds_tile=0, linear=0, fuse=0, **cfg)
    sim.step(40)
    orc.step(40)
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_linescan_matches_oracle(precision):
    """Device projection + host convolution (cs_linescan) against the
    oracle's direct sum of species.c:1960-2035, every axis, with and
    without the domain rule."""
    d = with_serca(dyad.small_dyad(seed=16))
    opts = cs.Options(seed=2, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0}, precision=precision)
    rest = Oracle(d.geometry, d.model, opts)     # t = 0: every field at rest
    sim, orc = _run_pair(d, opts, 60)
    for s, axis, offsets, nd in [(1, 0, (0, 2, -3), 0), (0, 1, (1, 0, 2), 2), (1, 2, (-2, 3, 0), 1),
                                 (2, 2, (0, 0, 0), 0), (0, 0, (0, -4, 1), 3)]:
        _say(f"linescan species {s} axis {axis} offsets {offsets} domains {nd}")
        got = sim.linescan(s, axis, offsets, d.model.conv, nd)
        want = orc.linescan(s, axis, offsets, d.model.conv, nd)
        if precision == "f64":
            np.testing.assert_allclose(got, want, rtol=RTOL_F64, atol=RTOL_F64 * float(np.max(np.abs(want))))
        else:
            # FP32 fields resolve the deviation from rest: so does their linescan
            ok, rel, err, dev = deviation_close(got, want, rest.linescan(s, axis, offsets, d.model.conv, nd), DEV_F32)
            _say(f"  fp32 error {err:.3e} = {rel:.2e} of the largest change {dev:.3e}")
            assert ok
    sim.close()
    orc.close()
    rest.close()


def test_linescan_wide_grid():
    """x scans over more lanes than a block has and y in several chunks."""
    d = dyad.small_dyad(n=(300, 70, 6), seed=22, num_ryr=16, num_pryr=4)
    opts = cs.Options(seed=3, open={"ryr": [d.open_index]})
    sim, orc = _run_pair(d, opts, 10)
    for axis, offsets in [(0, (0, 5, -1)), (1, (-20, 0, 1)), (2, (7, -3, 0))]:
        _say(f"wide linescan axis {axis}")
        got = sim.linescan(0, axis, offsets, d.model.conv)
        want = orc.linescan(0, axis, offsets, d.model.conv)
        np.testing.assert_allclose(got, want, rtol=RTOL_F64, atol=RTOL_F64 * float(np.max(np.abs(want))))
    sim.close()
    orc.close()


def test_autotune_keeps_state():
    d = dyad.small_dyad(seed=13)
    opts = cs.Options(seed=4, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    sim = cs.Simulation(d.geometry, d.model, opts)
    sim.step(10)
    before = [sim.species(s) for s in range(sim.ns)]
    cfg, ms = sim.autotune(reps=1)
    _say(f"autotune small: {cfg} {ms * 1e3:.1f} us")
    for s in range(sim.ns):
        np.testing.assert_array_equal(sim.species(s), before[s])
    orc = Oracle(d.geometry, d.model, opts)
    orc.step(10)
    sim.step(30)
    orc.step(30)
    _compare_fields(sim, orc, RTOL_F64)
    sim.close()


def test_stats_match_oracle():
    d = dyad.small_dyad(seed=2)
    opts = cs.Options(seed=5, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    sim, orc = _run_pair(d, opts, 50)
    a = sim.stats()
    b = orc.stats()
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-10, atol=0)
    sim.close()


def test_upload_download_roundtrip():
    d = dyad.small_dyad(seed=1)
    opts = cs.Options(seed=1)
    sim = cs.Simulation(d.geometry, d.model, opts)
    rng = np.random.default_rng(0)
    for s in range(sim.ns):
        x = rng.random(sim.n)
        sim.set_species(s, x)
        # the device stores x minus the domain's resting value r; the round
        # trip (x - r) + r is exact to within one rounding at the size of r
        r = float(np.max(np.abs(d.model.init[:, s])))
        np.testing.assert_allclose(sim.species(s), x, rtol=0, atol=2 * np.spacing(max(r, 1.0)))
    sim.close()


def test_cluster_dyad_conservation():
    """Full 400x400x200 grid (32M voxels): with no-flux faces and internal
    fluxes only, total calcium (free + bound, every species slot) is
    conserved; FP64 must keep it to round-off."""
    d = dyad.cluster_dyad(seed=1)
    opts = cs.Options(h=3.0, seed=1, open={"ryr": list(range(0, 60, 3))}, t_close={"ryr": 5.0})
    _say("case: 400x400x200 cluster dyad conservation fp64")
    sim = cs.Simulation(d.geometry, d.model, opts)
    assert sim.n == (400, 400, 200)
    tot0 = sim.stats()[2].sum()
    sim.step(50)
    mn, mx, tot = sim.stats()
    rel = abs(tot.sum() - tot0) / tot0
    _say(f"cluster dyad conservation rel drift {rel:.2e}")
    assert rel < 1e-12
    assert np.all(mn[:3] > -1e-9)
    sim.close()
