"""The 64x64x25 lattice dyad (384x384x150 nm at 6 nm, 49 RyRs of which 10
phosphorylated, the shipped-model kinetics) and fixed domain species on
the GPU against the oracle.

FP64 fields must match the oracle to 1e-10 per voxel in one-step and in
fused two-step passes.  FP32 fields are held per species to 1e-4 of that
species' largest deviation from its resting value: the quantity the
stored deviations resolve, and the one a spark changes.
"""
import dataclasses
import sys

import numpy as np
import pytest

import csparks as cs
import dyad
from helpers import DEV_F32, deviation_error, rest_field, species_close
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

RTOL_F64 = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lattice_oracle():
    """The lattice dyad with RyR 24 (the centre) open until 0.04 ms, 2000
    steps of the oracle (0.0518 ms), fields kept."""
    d = dyad.lattice_dyad(seed=1)
    opts = cs.Options(seed=1, open={"ryr": [d.open_index]}, t_close={"ryr": 0.04})
    _say(f"case: oracle {d.description} 2000 steps")
    orc = Oracle(d.geometry, d.model, opts)
    orc.step(1000)
    mid = [orc.species(s) for s in range(orc.ns)]
    orc.step(1000)
    end = [orc.species(s) for s in range(orc.ns)]
    states = [orc.states(k) for k in range(2)]
    orc.close()
    return d, opts, mid, end, states


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_lattice_dyad_fp64_matches_oracle(lattice_oracle, fuse):
    d, opts, mid, end, states = lattice_oracle
    _say(f"case: {d.description} fp64 fuse={fuse} 2000 steps")
    sim = cs.Simulation(d.geometry, d.model, opts)
    assert sim.n == (64, 64, 25)
    sim.set_launch_config(fuse=fuse)
    for want in (mid, end):
        sim.step(1000)
        for s in range(sim.ns):
            ok, err = species_close(sim.species(s), want[s], RTOL_F64)
            assert ok, f"species {s}: max rel err {err:.3e}"
    for k in range(2):
        np.testing.assert_array_equal(sim.channel_states(k), states[k])
    # the centre channel closed at 0.04 ms and took the spark with it
    assert not states[0].any()
    sim.close()


def test_lattice_dyad_fp32_error_against_deviation(lattice_oracle):
    """FP32 fields, the default fused passes, against the FP64 oracle."""
    d, opts, mid, end, states = lattice_oracle
    o32 = dataclasses.replace(opts, precision="f32")
    sim = cs.Simulation(d.geometry, d.model, o32)
    sim.step(2000)
    for s in range(sim.ns):
        err, dev = deviation_error(sim.species(s), end[s], rest_field(d, s))
        _say(f"species {s} ({d.model.species_names[s]}): max |u - rest| {dev:.4g} uM, "
             f"FP32 error {err:.3e} uM = {err / max(dev, 1e-300):.2e} of it")
        assert dev > 0
        assert err <= DEV_F32 * dev, f"species {s}: {err:.3e} > {DEV_F32} x {dev:.3e}"
    sim.close()


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_fixed_domain_species_fp64_matches_oracle(fuse):
    """(cleft, Fluo) and (jsr, CSQN slot) fixed: every kernel that writes
    those values (stencil, reaction, boundary flux) must leave them at rest."""
    d = dyad.small_dyad(seed=5)
    names = d.model.domain_names
    model = dataclasses.replace(d.model, fixed=[(names.index("cleft"), 1), (names.index("jsr"), 2)])
    opts = cs.Options(seed=2, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    _say(f"case: fixed domain species fuse={fuse}")
    orc = Oracle(d.geometry, model, opts)
    sim = cs.Simulation(d.geometry, model, opts)
    sim.set_launch_config(fuse=fuse)
    for _ in range(3):
        orc.step(47)
        sim.step(47)
        for s in range(sim.ns):
            ok, err = species_close(sim.species(s), orc.species(s), RTOL_F64)
            assert ok, f"species {s}: max rel err {err:.3e}"
    ids = d.geometry.voxels
    np.testing.assert_array_equal(sim.species(1)[ids == names.index("cleft")],
                                  np.asarray(model.init)[names.index("cleft"), 1])
    sim.close()
    orc.close()


@pytest.mark.parametrize("fuse", [0, 1], ids=["one-step", "fused"])
def test_fp32_small_dyad_error_against_deviation(fuse):
    """The small dyad with a channel open, FP32 fields, against the FP64
    oracle, per species relative to its largest deviation from rest."""
    d = dyad.small_dyad(n=(40, 36, 14), seed=6)
    opts = cs.Options(seed=5, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    orc = Oracle(d.geometry, d.model, opts)
    sim = cs.Simulation(d.geometry, d.model, dataclasses.replace(opts, precision="f32"))
    sim.set_launch_config(fuse=fuse)
    orc.step(401)
    sim.step(401)
    for s in range(sim.ns):
        err, dev = deviation_error(sim.species(s), orc.species(s), rest_field(d, s))
        _say(f"species {s}: dev {dev:.4g}, err {err:.3e} ({err / max(dev, 1e-300):.2e})")
        assert err <= DEV_F32 * dev, f"species {s}: {err:.3e} > {DEV_F32} x {dev:.3e}"
    sim.close()
    orc.close()
