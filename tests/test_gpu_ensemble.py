"""GPU parity of ensembles (cs_batch_*): many uneven realisations in one
launch, each against its own oracle run or its own single-domain context.

Deterministic and Mersenne-Twister FP64 runs must match the oracle to
1e-10 relative per voxel, with identical channel trajectories; Philox runs
must match a cs_context of the same problem (the same arithmetic, so the
same states and, in FP64, fields to 1e-12); FP32 fields are held to 1e-4
of each species' largest deviation from rest.  Captured-graph replays
must equal direct launches, retired realisations must keep their state,
and the 512-realisation workload must conserve total calcium per
realisation.
"""
import dataclasses
import sys

import numpy as np
import pytest

import csparks as cs
import dyad
from helpers import species_close
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

RTOL_F64 = 1e-10
SHAPES = [(20, 18, 14), (24, 20, 12), (16, 16, 14), (30, 22, 13), (18, 26, 15), (22, 22, 16), (27, 19, 14),
          (21, 30, 12)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _members(fast: bool, seed0: int = 11):
    """Eight small dyads of different shapes and cluster sizes (every one
    with both RyR and pRyR channels, so the flux lists agree); with fast
    set, RyR kinetics fast enough that channels flip within a few hundred
    steps."""
    out = []
    for k, n in enumerate(SHAPES):
        d = dyad.small_dyad(n=n, seed=seed0 + k, num_ryr=6 + (k % 4) * 3, num_pryr=2 + k % 3)
        if fast:
            kin = dict(Kd_open=5.0, k_max_open=200.0, k_min_close=50.0, k_max_close=50.0)
            m = d.model
            d = dyad.Dyad(d.geometry, dataclasses.replace(m, ryr=dataclasses.replace(m.ryr, **kin),
                                                           p_ryr=dataclasses.replace(m.p_ryr, **kin)),
                          d.open_index, d.description)
        out.append(d)
    return out


def _compare_member(ens, i, ref, rtol):
    worst = 0.0
    for s in range(ens.ns):
        ok, err = species_close(ens.species(i, s), ref.species(s), rtol)
        worst = max(worst, err)
        assert ok, f"member {i} species {s}: max rel err {err:.3e} > {rtol}"
    return worst


def test_ensemble_fp64_deterministic_matches_oracle():
    ds = _members(fast=False)
    opts = [cs.Options(seed=100 + k, open={"ryr": [d.open_index]}, t_close={"ryr": 0.003, "p_ryr": 0.002})
            for k, d in enumerate(ds)]
    _say(f"case: 8 deterministic members, shapes {SHAPES}")
    ens = cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)])
    orcs = [Oracle(d.geometry, d.model, o) for d, o in zip(ds, opts)]
    for chunk in (70, 90):
        ens.step(chunk)
        for o in orcs:
            o.step(chunk)
    worst = 0.0
    for i, o in enumerate(orcs):
        worst = max(worst, _compare_member(ens, i, o, RTOL_F64))
        for f in range(2):
            np.testing.assert_array_equal(ens.channel_states(i, f), o.states(f))
        o.close()
    _say(f"deterministic ensemble max rel err {worst:.2e}")
    ens.close()


def test_ensemble_fp64_mt_matches_oracle():
    """Stochastic gating from each member's own Mersenne Twister stream:
    identical channel trajectories and fields to 1e-10."""
    ds = _members(fast=True)
    opts = [cs.Options(seed=200 + k, open={"ryr": [d.open_index]}, rng="mt") for k, d in enumerate(ds)]
    _say("case: 8 stochastic (MT) members")
    ens = cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)])
    orcs = [Oracle(d.geometry, d.model, o) for d, o in zip(ds, opts)]
    flips = 0
    for chunk in (100, 100, 100):
        ens.step(chunk)
        for o in orcs:
            o.step(chunk)
        for i, o in enumerate(orcs):
            for f in range(2):
                np.testing.assert_array_equal(ens.channel_states(i, f), o.states(f), err_msg=f"member {i} flux {f}")
                flips += int(ens.recruited(i, f).sum())
    worst = max(_compare_member(ens, i, o, RTOL_F64) for i, o in enumerate(orcs))
    for i, o in enumerate(orcs):
        for f in range(2):
            np.testing.assert_array_equal(ens.recruited(i, f), o.recruited(f))
    assert flips > 8, "the fast kinetics should recruit channels beyond the forced-open ones"
    _say(f"MT ensemble max rel err {worst:.2e}, recruited (summed over checks) {flips}")
    ens.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_ensemble_equals_single_contexts_philox(precision):
    """A member takes the same steps as a cs_context of its problem: the
    same Philox key and counters, the same per-voxel arithmetic."""
    ds = _members(fast=True, seed0=31)
    opts = [cs.Options(seed=300 + k, open={"ryr": [d.open_index]}, precision=precision) for k, d in enumerate(ds)]
    _say(f"case: 8 Philox members vs single contexts, {precision}")
    ens = cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)], instances=list(range(8)))
    sims = []
    for k, (d, o) in enumerate(zip(ds, opts)):
        sim = cs.Simulation(d.geometry, d.model, o, instance=k)
        sim.set_launch_config(fuse=0)
        sims.append(sim)
    for chunk in (150, 150):
        ens.step(chunk)
        for s in sims:
            s.step(chunk)
    rtol = 1e-12 if precision == "f64" else 1e-6
    worst = 0.0
    for i, s in enumerate(sims):
        for f in range(2):
            np.testing.assert_array_equal(ens.channel_states(i, f), s.channel_states(f), err_msg=f"member {i}")
            np.testing.assert_array_equal(ens.recruited(i, f), s.recruited(f))
        worst = max(worst, _compare_member(ens, i, s, rtol))
        s.close()
    _say(f"ensemble vs contexts ({precision}) max rel err {worst:.2e}")
    ens.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("shape", [dict(rows=8, zchunk=8), dict(rows=1, zchunk=2)])
def test_tile_blocks_equal_persistent_blocks(precision, shape):
    """The stencil pass with one block per tile (persistent = 0, the tiles
    in groups per XCD once there are enough of them: 64-column tiles two
    planes deep make 446) and with persistent blocks taking tiles
    from the queues (persistent = 1) give the same fields and channel
    trajectories, bit for bit."""
    ds = _members(fast=True, seed0=61)
    opts = [cs.Options(seed=500 + k, open={"ryr": [d.open_index]}, precision=precision) for k, d in enumerate(ds)]
    _say(f"case: tiles vs persistent blocks, {precision}, {shape}")
    runs = []
    for persistent in (0, 1):
        e = cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)], instances=list(range(8)))
        e.set_launch(prefetch=1, nontemporal=0, persistent=persistent, **shape)
        runs.append(e)
    for chunk in (150, 41):
        for e in runs:
            e.step(chunk)
        for i in range(8):
            for f in range(2):
                np.testing.assert_array_equal(runs[0].channel_states(i, f), runs[1].channel_states(i, f))
            for s in range(runs[0].ns):
                np.testing.assert_array_equal(runs[0].species(i, s), runs[1].species(i, s),
                                              err_msg=f"member {i} species {s}")
    for e in runs:
        e.close()


def test_ensemble_graph_replays_equal_direct_launches():
    """Captured periods replayed several times, with eager reductions and
    synchronisations between the calls, equal direct launches."""
    ds = _members(fast=True, seed0=51)
    opts = [cs.Options(seed=400 + k, open={"ryr": [d.open_index]}) for k, d in enumerate(ds)]
    mk = lambda: cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)])
    a, b = mk(), mk()
    a.set_graph(True)
    for rep in range(4):
        n = 160 + 7 * rep          # whole periods plus a remainder
        a.step(n)
        b.step(n)
        ma, _, ta = a.stats()
        mb, _, tb = b.stats()
        np.testing.assert_array_equal(ta, tb)
        a.synchronize()
        for i in range(a.count):
            for f in range(2):
                np.testing.assert_array_equal(a.channel_states(i, f), b.channel_states(i, f))
            for s in range(a.ns):
                np.testing.assert_array_equal(a.species(i, s), b.species(i, s), err_msg=f"replay {rep} member {i}")
        _say(f"graph replay {rep}: equal after {a.time_ind - 1} steps")
    a.close()
    b.close()


def test_ensemble_retired_members_keep_their_state():
    ds = _members(fast=True, seed0=71)
    opts = [cs.Options(seed=500 + k, open={"ryr": [d.open_index]}) for k, d in enumerate(ds)]
    mk = lambda: cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)])
    a, b = mk(), mk()
    a.step(60)
    b.step(60)
    frozen = {i: [a.species(i, s) for s in range(a.ns)] for i in (1, 4, 6)}
    active = np.ones(a.count, dtype=bool)
    active[[1, 4, 6]] = False
    a.set_active(active)
    a.step(90)
    b.step(90)
    for i in range(a.count):
        for s in range(a.ns):
            if active[i]:
                np.testing.assert_array_equal(a.species(i, s), b.species(i, s))
            else:
                np.testing.assert_array_equal(a.species(i, s), frozen[i][s])
    a.close()
    b.close()


def test_ensemble_stats_and_linescans_match():
    ds = _members(fast=False, seed0=91)
    opts = [cs.Options(seed=600 + k, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0, "p_ryr": 1.0})
            for k, d in enumerate(ds)]
    ens = cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)])
    orcs = [Oracle(d.geometry, d.model, o) for d, o in zip(ds, opts)]
    ens.step(120)
    for o in orcs:
        o.step(120)
    mn, mx, tot = ens.stats()
    conv = ds[0].model.conv
    scans = ens.linescans(1, 0, (0, 0, -3), conv, num_domains=2)
    for i, o in enumerate(orcs):
        omn, omx, otot = o.stats()
        np.testing.assert_allclose(mn[i], omn, rtol=1e-10, atol=0)
        np.testing.assert_allclose(mx[i], omx, rtol=1e-10, atol=0)
        np.testing.assert_allclose(tot[i], otot, rtol=1e-10, atol=1e-300)
        np.testing.assert_allclose(scans[i], o.linescan(1, 0, (0, 0, -3), conv, 2), rtol=1e-10, atol=0)
        o.close()
    ens.close()


def _total_calcium(tot):
    """Free plus buffer-bound calcium per member: every species slot of the
    dyad model holds one Ca per molecule (Ca, Ca-Fluo, Ca-CMDN/CSQN,
    Ca-ATP, Ca-TRPN)."""
    return tot.sum(axis=(1, 2))


def test_config2_ensemble_conserves_calcium_and_fp32_tracks_fp64():
    """512 realisations of the ensemble workload (edges U[256, 512] nm at
    6 nm, 9-100 RyRs, 0-50% pRyR, one trigger RyR open, Philox gating):
    FP64 conserves total calcium per realisation to 1e-12 relative (RyR
    fluxes move calcium between two voxels, reactions between slots, the
    outer faces are closed); FP32 fields stay within 1e-4 of each
    species' largest deviation from rest."""
    ds = [dyad.ensemble_dyad(seed) for seed in range(512)]
    _say(f"case: 512 ensemble dyads, {sum(d.geometry.voxels.size for d in ds)} voxels")
    res = {}
    for prec in ("f64", "f32"):
        opts = [cs.Options(seed=1000 + k, open={"ryr": [d.open_index]}, precision=prec) for k, d in enumerate(ds)]
        ens = cs.Ensemble([(d.geometry, d.model, o) for d, o in zip(ds, opts)], instances=list(range(512)))
        t0 = _total_calcium(ens.stats()[2])
        ens.step(400)
        t1 = _total_calcium(ens.stats()[2])
        drift = np.abs(t1 - t0) / t0
        _say(f"{prec}: max relative drift of total calcium {drift.max():.2e}")
        if prec == "f64":
            assert drift.max() < 1e-12
        res[prec] = ens
    worst = 0.0
    for i in (0, 77, 300, 511):
        for s in range(res["f64"].ns):
            a, b = res["f32"].species(i, s), res["f64"].species(i, s)
            rest = ds[i].model.init[ds[i].geometry.voxels, s]
            dev = float(np.max(np.abs(b - rest)))
            err = float(np.max(np.abs(a - b)))
            if dev > 0:
                worst = max(worst, err / dev)
                assert err <= 1e-4 * dev, f"member {i} species {s}: {err:.3e} vs deviation {dev:.3e}"
    _say(f"fp32 vs fp64: max error / max deviation {worst:.2e}")
    for e in res.values():
        e.close()
