"""Host logic of the closure's training and judging (no GPU): the spark
kinetics of closure_sparks, the linescan offsets on both grids, the
release ratio of the flux network's labels, and the normalisation."""
import math

import numpy as np
import pytest

import closure
import closure_sparks
import csparks as cs
import dyad


def test_spark_model_speeds_up_gating_and_keeps_the_pryr_sensitivity_higher():
    d = dyad.ensemble_dyad(3, h=12.0, edge_nm=(192.0, 256.0), ryr=(9, 30))
    m = closure_sparks.spark_model(d.model)
    assert m.ryr.k_max_open == 200.0 and m.ryr.k_min_close == m.ryr.k_max_close == 50.0
    assert m.p_ryr.Kd_open == m.ryr.Kd_open / 2
    assert m.ryr.n_open == d.model.ryr.n_open            # the power stays the shipped one
    assert m.ryr.i_unitary == d.model.ryr.i_unitary


def test_line_offsets_point_at_the_same_plane_on_both_grids():
    d = dyad.ensemble_dyad(4, h=12.0, edge_nm=(192.0, 256.0), ryr=(9, 30))
    nz = d.geometry.voxels.shape[2]
    zc = int(np.argwhere(d.geometry.voxels == dyad.CLEFT)[:, 2].mean())
    o12 = closure_sparks.line_offsets(d, 12.0)
    o3 = closure_sparks.line_offsets(d, 3.0)
    assert o12[:2] == (0, 0) and o3[:2] == (0, 0)
    # the line's height above the bottom face: 12 nm voxel centre vs 3 nm
    z12 = (nz // 2 + o12[2] + 0.5) * 12.0
    z3 = (4 * nz // 2 + o3[2] + 0.5) * 3.0
    assert abs(z12 - (zc + 0.5) * 12.0) < 1e-9
    assert abs(z3 - z12) <= 3.0


def test_release_ratio_in_the_linear_regime_is_the_ratio_of_differences():
    # k = g dt / V: the fine factor over the coarse one is 16 steps x 4^3 volume -> k_f * 16 / (k_c * 64) = 1
    k_c = 0.01
    k_f = k_c * 64 / 16
    r = closure.release_ratio([1000.0, 500.0], [100.0, 400.0], [1300.0, 1300.0], [100.0, 100.0], k_f, k_c, 16, 4.0)
    np.testing.assert_allclose(r, [900.0 / 1200.0, 100.0 / 1200.0], rtol=1e-14)
    # no release when the fine mouth has equilibrated or reversed
    r = closure.release_ratio([300.0], [400.0], [1300.0], [100.0], k_f, k_c, 16, 4.0)
    assert r[0] == 0.0


def test_ryr_release_factor_is_the_two_voxel_relaxation():
    """(1 - K0)/2 with K0 = exp(-2 dt 1e15 / (2 F) i / 1299.9 / h^3)
    (ryr_flux, boundaryfluxes.c:87-141), at the 3 and 12 nm dts."""
    lib = cs.device_lib()
    p = np.zeros(10)
    p[6] = 0.194985
    ptr = p.ctypes.data_as(cs.C.POINTER(cs.C.c_double))
    for h, dt in ((3.0, 6.4772727272727265e-06), (12.0, 1.0363636363636362e-04)):
        k0 = math.exp(-2 * dt * 1e15 / (2 * 96485.0) * p[6] / 1299.9 / h ** 3)
        assert lib.cs_ryr_flux(dt, h, 1.0, 0.0, ptr) == pytest.approx((1 - k0) / 2, rel=1e-12)


def test_normalisation_round_trips_and_flags_nothing_inside_its_range():
    rng = np.random.default_rng(0)
    raw = np.abs(rng.normal(10.0, 3.0, (500, cs.MLP_IN))).astype(np.float32)
    ca = np.abs(rng.normal(50.0, 20.0, 500)) + 1.0
    norm = closure.fit_normalisation(raw, ca)
    x = norm.inputs(raw)
    assert np.all(x >= norm.in_lo - 1e-5) and np.all(x <= norm.in_hi + 1e-5)
    np.testing.assert_allclose(norm.concentration(norm.target(ca)), ca, rtol=1e-5)
