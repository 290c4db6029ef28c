"""GPU: fused two-step passes (k_step2 with k_bridge) against one-step
passes and against the CPU oracle.

A fused pass computes step t+1 from step t without writing step t to HBM,
and applies the fluxes and gating of both steps around it; the result is
the same sequence of per-voxel updates as two one-step passes.  FP64 runs
are held to the oracle at 1e-10 and to the one-step path at 1e-13, with
identical channel trajectories; FP32 runs to the one-step path within a
few FP32 roundings of the field's deviation from rest.  Grids cover tile
edges in x (62 inner columns per tile; nx = 63 and 125 put the last
column on a tile's last lane, whose face neighbour is its own column) and
y (14 rows in FP32, 6 in FP64),
one- and two-plane grids, z chunks that do not divide the grid, odd step
counts (a one-step pass at the end) and runs split into several calls.
"""
import sys

import numpy as np
import pytest

import csparks as cs
import dyad
import sparks
from helpers import (DEV_F32, deviation_close, load_geometry_fixture, load_model_fixture, rest_field, species_close,
                     with_serca)
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

RTOL_F64 = 1e-10
DEV_FUSED_F32 = 1e-5   # fused against one-step FP32 passes, of each species' largest deviation from rest


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fused(sim, zchunk=8, xcd=0):
    ok, why = sim.fused_available()
    assert ok, why
    sim.set_launch_config(fuse=1, fuse_zchunk=zchunk, fuse_xcd=xcd)


def _run(d, opts, chunks, zchunk=8, xcd=0, oracle=True):
    """Fused run, one-step run and (optionally) the oracle over the same
    step chunks."""
    _say(f"case: {d.description} n={d.geometry.voxels.shape} opts={opts} chunks={chunks} "
         f"fuse_zchunk={zchunk} fuse_xcd={xcd}")
    fu = cs.Simulation(d.geometry, d.model, opts)
    _fused(fu, zchunk, xcd)
    one = cs.Simulation(d.geometry, d.model, opts)
    one.set_launch_config(fuse=0)
    orc = Oracle(d.geometry, d.model, opts) if oracle else None
    for k in chunks:
        fu.step(k)
        one.step(k)
        if orc:
            orc.step(k)
        for f in range(sparks.discrete_count(fu)):
            np.testing.assert_array_equal(fu.channel_states(f), one.channel_states(f))
            if orc:
                np.testing.assert_array_equal(fu.channel_states(f), orc.states(f))
    fu.synchronize()
    return fu, one, orc


def _close_dev(fu, other, d, rtol, what):
    """Per species: |fu - other| <= rtol x max |other - rest|."""
    worst = 0.0
    for s in range(fu.ns):
        ok, rel, err, dev = deviation_close(fu.species(s), other(s), rest_field(d, s), rtol)
        _say(f"species {s} against {what}: error {err:.3e} = {rel:.2e} of the largest deviation {dev:.3e}")
        worst = max(worst, rel)
        assert ok, f"species {s} against {what}: {err:.3e} > {rtol} x {dev:.3e}"
    return worst


def _close(fu, other, rtol, what):
    worst = 0.0
    for s in range(fu.ns):
        ok, err = species_close(fu.species(s), other(s), rtol)
        worst = max(worst, err)
        assert ok, f"species {s} against {what}: max rel err {err:.3e} > {rtol}"
    return worst


@pytest.mark.parametrize("n,zchunk,xcd", [((20, 18, 14), 5, 0), ((65, 13, 9), 4, 1), ((130, 30, 6), 64, 8),
                                          ((61, 16, 5), 1, 0), ((9, 9, 5), 3, 2), ((63, 12, 7), 2, 0),
                                          ((125, 9, 6), 7, 1)])
def test_fused_fp64_deterministic_matches_oracle(n, zchunk, xcd):
    d = dyad.small_dyad(n=n, seed=3)
    opts = cs.Options(seed=7, open={"ryr": [d.open_index, 2]}, t_close={"ryr": 1.0})
    fu, one, orc = _run(d, opts, [41, 40, 3], zchunk, xcd)
    w1 = _close(fu, orc.species, RTOL_F64, "oracle")
    w2 = _close(fu, one.species, 1e-13, "one-step passes")
    _say(f"max rel err: oracle {w1:.2e}, one-step {w2:.2e}")
    assert fu.time_ind == one.time_ind
    fu.close()
    one.close()


@pytest.mark.parametrize("nz,zchunk", [(1, 1), (2, 1), (2, 5)])
def test_fused_thin_slabs(nz, zchunk):
    """One- and two-plane grids (a slab of a dyad, no channels): the z
    neighbours clamp at both faces on both levels."""
    import dataclasses
    d = dyad.small_dyad(n=(61, 16, 8), seed=3)
    g = d.geometry
    geom = cs.Geometry(voxels=np.ascontiguousarray(g.voxels[:, :, 3:3 + nz]), h=g.h,
                       domain_names=list(g.domain_names), domain_connections=g.domain_connections, boundaries=[])
    slab = dyad.Dyad(geom, dataclasses.replace(d.model, ryr=None, p_ryr=None), 0, f"slab of {nz} planes")
    opts = cs.Options(seed=1)
    fu, one, orc = _run(slab, opts, [0], zchunk, 0)
    # a perturbed resting state, the same on all three
    x = [orc.species(s) * (1 + 0.2 * np.random.default_rng(s).random(orc.species(s).shape)) for s in range(orc.ns)]
    for sim in (fu, one, orc):
        for s in range(orc.ns):
            sim.set_species(s, x[s])
    for k in (31, 30):
        fu.step(k)
        one.step(k)
        orc.step(k)
    _close(fu, orc.species, RTOL_F64, "oracle")
    _close(fu, one.species, 1e-13, "one-step passes")


@pytest.mark.parametrize("n,precision", [((63, 12, 7), "f64"), ((125, 9, 6), "f64"), ((63, 17, 8), "f32")])
def test_fused_face_on_last_lane(n, precision):
    """nx = 62k + 1 puts the grid's last column on a tile's last lane, whose
    x neighbour across the face is its own value on every plane.  The fields
    start perturbed by 20%, so the faces carry gradients from the first
    step and a stale or wrong neighbour there shows at once."""
    d = dyad.small_dyad(n=n, seed=3)
    opts = cs.Options(seed=7, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0}, precision=precision)
    fu, one, orc = _run(d, opts, [0], zchunk=3, xcd=0)
    x = [orc.species(s) * (1 + 0.2 * np.random.default_rng(s).random(orc.species(s).shape)) for s in range(orc.ns)]
    for sim in (fu, one, orc):
        for s in range(orc.ns):
            sim.set_species(s, x[s])
    for k in (30, 31):
        fu.step(k)
        one.step(k)
        orc.step(k)
    if precision == "f64":
        _close(fu, orc.species, RTOL_F64, "oracle")
        _close(fu, one.species, 1e-13, "one-step passes")
    else:
        # the fields start 20% off rest everywhere, so the deviation is large
        _close_dev(fu, orc.species, d, DEV_F32, "oracle")
        _close_dev(fu, one.species, d, DEV_FUSED_F32, "one-step passes")


def test_fused_forced_close_and_substeps():
    """-C closes the forced-open channel mid-run; reaction and gating
    substeps > 1, so some pairs gate or react on one step only."""
    d = dyad.small_dyad(seed=5)
    opts = cs.Options(seed=3, open={"ryr": [d.open_index]}, t_close={"ryr": 0.002, "p_ryr": 0.001},
                      dt_update_react=3.5e-5 * 3, dt_update_stoch=3.5e-5 * 2)
    fu, one, orc = _run(d, opts, [121, 119])
    assert fu.problem.reaction_substep > 1 and fu.problem.stochastic_substep > 1
    _close(fu, orc.species, RTOL_F64, "oracle")
    _close(fu, one.species, 1e-13, "one-step passes")
    assert fu.channel_states(0).sum() == 0


def test_fused_stochastic_philox_same_trajectories():
    """Stochastic gating with Philox: fused and one-step runs draw the
    same uniforms and see the same Ca, so every channel trajectory agrees
    (checked after every chunk), and so do the fields."""
    d = dyad.fast_spark_dyad()
    opts = cs.Options(seed=11, open={"ryr": [d.open_index]})
    fu, one, _ = _run(d, opts, [50] * 8, zchunk=6, xcd=1, oracle=False)
    rec_f = [fu.recruited(f) for f in range(sparks.discrete_count(fu))]
    rec_o = [one.recruited(f) for f in range(sparks.discrete_count(one))]
    for a, b in zip(rec_f, rec_o):
        np.testing.assert_array_equal(a, b)
    _say(f"recruited {sum(int(a.sum()) for a in rec_f)}")
    _close(fu, one.species, 1e-13, "one-step passes")


def test_fused_stochastic_mt_matches_oracle():
    """The reference's Mersenne Twister draw order, uniforms from the host:
    a fused pass consumes them in step order."""
    d = dyad.fast_spark_dyad()
    opts = cs.Options(seed=4, rng="mt", open={"ryr": [d.open_index]})
    fu, one, orc = _run(d, opts, [60, 61, 60])
    _close(fu, orc.species, RTOL_F64, "oracle")
    _close(fu, one.species, 1e-13, "one-step passes")


def test_fused_serca_and_dirichlet():
    d = with_serca(dyad.small_dyad(n=(65, 14, 10), seed=9))
    opts = cs.Options(seed=2, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0}, force_dirichlet=True)
    fu, one, orc = _run(d, opts, [60, 41], zchunk=3, xcd=0)
    _close(fu, orc.species, RTOL_F64, "oracle")
    _close(fu, one.species, 1e-13, "one-step passes")


@pytest.mark.parametrize("n,zchunk,xcd", [((130, 30, 12), 5, 8), ((63, 17, 9), 4, 0)])
def test_fused_fp32_matches_one_step(n, zchunk, xcd):
    """FP32: the fused pass applies the same per-voxel updates in the same
    order, so it agrees with one-step passes to FP32 rounding of the
    stored deviations."""
    d = dyad.small_dyad(n=n, seed=6)
    opts = cs.Options(seed=5, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0}, precision="f32")
    fu, one, orc = _run(d, opts, [100, 101], zchunk=zchunk, xcd=xcd)
    _close_dev(fu, orc.species, d, DEV_F32, "oracle")
    _close_dev(fu, one.species, d, DEV_FUSED_F32, "one-step passes")


def test_fused_full_dyad_fp32_conserves_and_matches():
    """The 400x400x200 dyad with the library's default launch shape (fused
    passes, its own z chunk and XCD groups): 20 steps agree with one-step
    passes in every species slot to DEV_FUSED_F32 of that species' largest
    deviation from rest, and with closed faces both runs keep total
    calcium (free plus bound in every slot) at its initial value."""
    d = dyad.cluster_dyad(seed=1)
    opts = cs.Options(h=3.0, seed=1, open={"ryr": [0, 1, 2]}, t_close={"ryr": 1.0}, precision="f32")
    fu = cs.Simulation(d.geometry, d.model, opts)
    cfg = fu.launch_config()
    _say(f"case: {d.description} default launch {cfg}")
    assert cfg["fuse"] == 1
    one = cs.Simulation(d.geometry, d.model, opts)
    one.set_launch_config(fuse=0)
    tot0 = fu.stats()[2].sum()
    fu.step(20)
    one.step(20)
    for f in range(sparks.discrete_count(fu)):
        np.testing.assert_array_equal(fu.channel_states(f), one.channel_states(f))
    _close_dev(fu, one.species, d, DEV_FUSED_F32, "one-step passes")
    for sim, what in ((fu, "fused"), (one, "one-step")):
        drift = abs(sim.stats()[2].sum() - tot0) / tot0
        _say(f"{what}: relative drift of total calcium {drift:.2e}")
        assert drift <= 1e-6
    fu.close()
    one.close()


def test_fuse_flag_checked():
    d = dyad.small_dyad()
    sim = cs.Simulation(d.geometry, d.model, cs.Options(seed=1))
    with pytest.raises(RuntimeError):
        sim.set_launch_config(fuse=2)
    with pytest.raises(RuntimeError):
        sim.set_launch_config(fuse=1, fuse_zchunk=0)
    sim.close()


def test_default_form_follows_the_fused_grid_size():
    """cs_create runs fused passes by default in FP32 only when the fused
    launch has at least half as many blocks as the device has CUs: the
    84^3 G1 geometry at 12 nm (36 blocks) takes one-step passes, the
    200x200x100 dyad (240 blocks) fused ones; FP64 always one-step."""
    g, m = load_geometry_fixture("g1_nonP_0"), load_model_fixture("params_sr1100_ryr55")
    d = dyad.cluster_dyad(seed=1, n=(200, 200, 100))
    cases = [(g, m, 12.0, "f32", 0), (d.geometry, d.model, 3.0, "f32", 1), (d.geometry, d.model, 3.0, "f64", 0)]
    for geo, model, h, prec, want in cases:
        sim = cs.Simulation(geo, model, cs.Options(h=h, seed=1, open={"ryr": [0]}, precision=prec))
        ok, why = sim.fused_available()
        cfg = sim.launch_config()
        _say(f"case: {sim.n} {prec}: fused available {ok}, default fuse {cfg['fuse']}")
        assert ok, why
        assert cfg["fuse"] == want
        sim.close()
