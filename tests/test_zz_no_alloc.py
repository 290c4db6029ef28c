"""cs_linescan_noalloc against cs_linescan, through the C ABI: the same
projection (to 1e-10 relative, the FP bar of the path) on two calls in a
row, the second on a smaller grid, for every axis; sizes beyond the fixed
scratch are rejected."""
import ctypes as C
import sys

import numpy as np
import pytest

import csparks as cs
import dyad

pytestmark = pytest.mark.gpu


def _weights(sim, conv, offsets):
    n, h = sim.n, sim.setup.problem.h
    off = [(n[ax] // 2 + int(offsets[ax])) * h for ax in range(3)]
    return [np.ascontiguousarray(cs.gauss(off[ax] - np.arange(n[ax]) * h, conv[ax])) for ax in range(3)]


def _both(sim, species, axis, w, mask):
    lib = cs.device_lib()
    lib.cs_linescan_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32] + [C.POINTER(C.c_double)] * 3 + \
        [C.c_uint32, C.POINTER(C.c_double)]
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    k = sim.n[axis]
    a, b = np.full(k, np.nan), np.full(k, np.nan)
    cs._check(lib.cs_linescan(sim.ctx, species, axis, ptr(w[0]), ptr(w[1]), ptr(w[2]), mask, ptr(a)), "cs_linescan")
    rc = lib.cs_linescan_noalloc(sim.ctx, species, axis, ptr(w[0]), ptr(w[1]), ptr(w[2]), mask, ptr(b))
    cs._check(rc, "cs_linescan_noalloc")
    return a, b


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_noalloc_linescan_equals_linescan_on_two_calls(precision):
    big = dyad.small_dyad(n=(70, 66, 20), seed=4)
    small = dyad.small_dyad(n=(20, 18, 14), seed=2)
    for d, steps in ((big, 80), (small, 40)):          # second call: smaller input
        print(f"case: {d.description} {precision}", flush=True)
        sim = cs.Simulation(d.geometry, d.model, cs.Options(seed=5, open={"ryr": [d.open_index]},
                                                            precision=precision))
        sim.step(steps)
        w = _weights(sim, d.model.conv, (0, 0, -3))
        for axis in (0, 1, 2):
            for species in (0, 1):
                a, b = _both(sim, species, axis, w, 0x3)
                assert np.all(np.isfinite(b))
                np.testing.assert_allclose(b, a, rtol=1e-10, atol=0)
        sim.close()


def test_noalloc_linescan_rejects_grids_beyond_its_scratch():
    """A 1 x 1100 x 1000 slab needs ny*nz = 1.1M partial sums on axes 1 and
    2, beyond the fixed scratch; axis 0 fits."""
    nx, ny, nz = 1, 1100, 1000
    m = dyad.dyad_species_model(with_pryr=False)
    m = cs.SpeciesModel(domain_names=m.domain_names, species_names=m.species_names, diffusive=m.diffusive,
                        sigma=m.sigma, init=m.init, buffers=m.buffers)
    g = cs.Geometry(voxels=np.zeros((nx, ny, nz), np.uint8), h=6.0, domain_names=list(dyad.DOMAINS),
                    domain_connections=dyad.connections())
    sim = cs.Simulation(g, m, cs.Options(seed=1, precision="f32"))
    w = _weights(sim, m.conv, (0, 0, 0))
    lib = cs.device_lib()
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    out = np.zeros(max(sim.n))
    lib.cs_linescan_noalloc.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32] + [C.POINTER(C.c_double)] * 3 + \
        [C.c_uint32, C.POINTER(C.c_double)]
    assert lib.cs_linescan_noalloc(sim.ctx, 0, 1, ptr(w[0]), ptr(w[1]), ptr(w[2]), 0x1F, ptr(out)) == -4   # CS_ERANGE
    assert b"scratch" in lib.cs_last_error()
    a, b = _both(sim, 0, 0, w, 0x1F)
    np.testing.assert_allclose(b, a, rtol=1e-10, atol=0)
    sim.close()
    sys.stdout.flush()
