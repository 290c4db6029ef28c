"""Pins for the CPU oracle (CPU only, no GPU).

The reference ships no tests or golden vectors (its README's test folder is
absent), so the oracle is pinned by
  * the reference's own functions, compiled from its sources into
    oracle/_ref (kernel_generic.cpp: kernel_SIMD, init_stencil_test2,
    reaction_SIMD; boundaryfluxes.c: ryr_flux, serca_flux,
    init_RyR_model_states_stochastically, evaluate_RyR_stochastically;
    mtwist.c: mt_seed32new, mt_ldrand),
  * published known answers (MT19937, Philox4x32-10),
  * closed forms and invariants of the scheme (constant fields, mass
    conservation, Neumann eigenmodes, interface coefficients, buffer
    conservation, two-voxel relaxation).
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import csparks as cs
import dyad
import oracle_py as op
from conftest import REF_BUILT

REF_K = os.path.join(REF_BUILT, "libref_kernels.so")
REF_F = os.path.join(REF_BUILT, "libref_fluxes.so")
need_ref_k = pytest.mark.skipif(not os.path.exists(REF_K), reason="oracle/_ref not built (needs the reference tree)")
need_ref_f = pytest.mark.skipif(not os.path.exists(REF_F), reason="oracle/_ref not built (needs the reference tree)")

dp = C.POINTER(C.c_double)


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


@pytest.fixture(scope="module")
def reff():
    L = C.CDLL(REF_F)
    L.ryr_flux.argtypes = [C.c_double] * 4 + [dp]
    L.ryr_flux.restype = C.c_double
    L.serca_flux.argtypes = [C.c_double] * 4 + [dp]
    L.serca_flux.restype = C.c_double
    L.mt_seed32new.argtypes = [C.c_uint32]
    L.mt_seed32new.restype = None
    L.mt_ldrand.restype = C.c_double
    L.mt_lrand.restype = C.c_uint32
    return L


@pytest.fixture(scope="module")
def refk():
    L = C.CDLL(REF_K)
    return L


def ryr_params(i_c=0.194985, Kd_open=107.6, n_open=2.8, k_max_open=0.3, k_min_open=1e-6,
               Kd_close=35.0, n_close=-0.5, k_min_close=0.9, k_max_close=3.0, t_close=-1.0):
    return np.array([Kd_close, k_max_close, n_open, k_max_open, k_min_open, n_close, i_c,
                     k_min_close, t_close, Kd_open], dtype=np.float64)


# ------------------------------------------------------------ generators
def test_mt19937_known_answers():
    """init_genrand(5489): first output 3499211612 and 10000th output
    4123659995 (the value C++ [rand.predef] requires of mt19937)."""
    h = cs.host_lib()
    g = cs.CMT()
    h.cs_mt_seed(C.byref(g), 5489)
    first = h.cs_mt_u32(C.byref(g))
    for _ in range(9998):
        h.cs_mt_u32(C.byref(g))
    assert first == 3499211612
    assert h.cs_mt_u32(C.byref(g)) == 4123659995
    L = op.lib()
    L.oracle_mt_seed(5489)
    assert L.oracle_mt_u32() == 3499211612
    for _ in range(9998):
        L.oracle_mt_u32()
    assert L.oracle_mt_u32() == 4123659995


@need_ref_f
@pytest.mark.parametrize("seed", [1, 7, 123456789])
def test_mt_ldrand_matches_reference(reff, seed):
    """2000 draws of mt_ldrand after mt_seed32new(seed): reference mtwist,
    the product's generator and the oracle's must be identical."""
    reff.mt_seed32new(seed)
    ref = np.array([reff.mt_ldrand() for _ in range(2000)])
    h = cs.host_lib()
    g = cs.CMT()
    h.cs_mt_seed(C.byref(g), seed)
    mine = np.array([h.cs_mt_ldrand(C.byref(g)) for _ in range(2000)])
    L = op.lib()
    L.oracle_mt_seed(seed)
    orc = np.array([L.oracle_mt_ldrand() for _ in range(2000)])
    np.testing.assert_array_equal(mine, ref)
    np.testing.assert_array_equal(orc, ref)


PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    """Random123 known-answer vectors for philox4x32-10."""
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o1 = (C.c_uint32 * 4)()
    o2 = (C.c_uint32 * 4)()
    op.lib().oracle_philox(c, k, o1)
    cs.device_lib().cs_philox4x32_10(c, k, o2)
    assert tuple(o1) == expect
    assert tuple(o2) == expect


# ------------------------------------------------------------ flux models
@need_ref_f
def test_ryr_flux_matches_reference(reff):
    rng = np.random.default_rng(0)
    for _ in range(500):
        dt, h = 10 ** rng.uniform(-6, -3), rng.uniform(3, 12)
        u0, u1 = rng.uniform(0, 1500), rng.uniform(0, 200)
        p = ryr_params(i_c=rng.uniform(0.05, 0.6))
        ref = reff.ryr_flux(dt, h, u0, u1, _ptr(p))
        assert op.lib().oracle_ryr_flux(dt, h, u0, u1, _ptr(p)) == pytest.approx(ref, rel=1e-15, abs=0)
        assert cs.device_lib().cs_ryr_flux(dt, h, u0, u1, _ptr(p)) == pytest.approx(ref, rel=1e-13, abs=0)


def test_ryr_flux_two_voxel_relaxation():
    """J conserves u0+u1 and relaxes u0-u1 by K0 = exp(-2 dt g), the exact
    solution of d(u0-u1)/dt = -2 g (u0-u1) with g = 1e15 i_c / (2 F 1299.9 h^3)."""
    p = ryr_params(i_c=0.2)
    dt, h, u0, u1 = 2e-5, 6.0, 1300.0, 0.1
    J = op.lib().oracle_ryr_flux(dt, h, u0, u1, _ptr(p))
    g = 1e15 * 0.2 / (2 * 96485.0 * 1299.9 * h ** 3)
    a0, a1 = u0 - J, u1 + J
    assert a0 + a1 == pytest.approx(u0 + u1, rel=1e-15)
    assert (a0 - a1) == pytest.approx((u0 - u1) * math.exp(-2 * dt * g), rel=1e-12)


@need_ref_f
def test_serca_flux_matches_reference(reff):
    rng = np.random.default_rng(1)
    for _ in range(500):
        dt, h = 10 ** rng.uniform(-6, -3), rng.uniform(3, 12)
        u0, u1 = 10 ** rng.uniform(-2, 2), rng.uniform(100, 1500)
        p = np.array([307.0, rng.uniform(1e-3, 100), 1.0])
        ref = reff.serca_flux(dt, h, u0, u1, _ptr(p))
        assert op.lib().oracle_serca_flux(dt, h, u0, u1, _ptr(p)) == pytest.approx(ref, rel=1e-14, abs=0)
        assert cs.device_lib().cs_serca_flux(dt, h, u0, u1, _ptr(p)) == pytest.approx(ref, rel=1e-12, abs=0)


# ------------------------------------------------------------- stencil
def _random_domains(rng, shape, nd):
    ids = np.zeros(shape, dtype=np.uint8)
    # blocky domains so most voxels sit in uniform neighbourhoods
    for _ in range(6):
        d = rng.integers(1, nd)
        lo = [rng.integers(0, s) for s in shape]
        hi = [min(s, l + rng.integers(2, 6)) for l, s in zip(lo, shape)]
        ids[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = d
    return ids


def _coef_table(rng, nd, conn):
    sig = rng.uniform(0.05, 0.15, nd)
    C5 = np.zeros((5, 5))
    for a in range(nd):
        C5[a, a] = sig[a]
        for b in range(a):
            if conn[a, b]:
                C5[a, b] = C5[b, a] = 0.5 * (sig[a] + sig[b])
    return C5


@need_ref_k
@pytest.mark.parametrize("shape", [(6, 7, 19), (5, 5, 24), (8, 4, 9), (4, 6, 33)])
def test_interior_stencil_matches_kernel_simd(refk, shape):
    """oracle stencil == reference kernel_SIMD (AVX-512 fast and general
    paths) on the interior x in [1,nx-2], y in [1,ny-2], all z."""
    rng = np.random.default_rng(sum(shape))
    nd = 3
    conn = np.ones((5, 5), dtype=bool)
    conn[2, 0] = conn[0, 2] = False
    C5 = _coef_table(rng, nd, conn)
    ids = _random_domains(rng, shape, nd)
    u = rng.uniform(0.1, 10.0, shape)
    pad = 128
    u_buf = np.zeros(u.size + pad)
    u_buf[:u.size] = u.ravel()
    du_buf = np.zeros(u.size + pad)
    id_buf = np.zeros(ids.size + pad, dtype=np.uint8)
    id_buf[:ids.size] = ids.ravel()
    test_buf = np.zeros(ids.size + pad, dtype=np.int8)
    nx, ny, nz = shape
    refk.init_stencil_test2(_ptr(id_buf, C.c_uint8), _ptr(test_buf, C.c_int8), nx, ny, nz)
    ghosts = (C.c_void_p * 6)()
    coef = np.ascontiguousarray(C5.ravel())
    refk.kernel_SIMD(_ptr(u_buf), _ptr(du_buf), _ptr(coef), _ptr(id_buf, C.c_uint8), ghosts,
                     _ptr(test_buf, C.c_int8), nx, ny, nz, C.c_ubyte(0), C.c_ubyte(0))
    ref = du_buf[:u.size].reshape(shape)
    mine = op.stencil(u, C5, ids)
    np.testing.assert_allclose(mine[1:-1, 1:-1, :], ref[1:-1, 1:-1, :], rtol=1e-14, atol=0)


def test_stencil_constant_field_is_steady():
    rng = np.random.default_rng(2)
    shape = (7, 6, 5)
    ids = _random_domains(rng, shape, 3)
    C5 = _coef_table(rng, 3, np.ones((5, 5), bool))
    u = np.full(shape, 3.25)
    np.testing.assert_array_equal(op.stencil(u, C5, ids), u)


def test_stencil_conserves_mass_with_noflux_faces():
    rng = np.random.default_rng(3)
    shape = (9, 7, 8)
    ids = _random_domains(rng, shape, 4)
    C5 = _coef_table(rng, 4, rng.random((5, 5)) > 0.3)
    C5 = np.minimum(C5, C5.T)  # symmetric like the reference table
    u = rng.uniform(0, 5, shape)
    assert op.stencil(u, C5, ids).sum() == pytest.approx(u.sum(), rel=1e-14)


@pytest.mark.parametrize("axis,k", [(0, 1), (1, 3), (2, 2), (0, 5)])
def test_stencil_neumann_eigenmode(axis, k):
    """cos(pi k (i+1/2)/N) is an eigenvector of the no-flux 7-point operator:
    one step multiplies it by 1 - 4 c sin^2(pi k / (2N)), faces included."""
    shape = (12, 10, 9)
    N = shape[axis]
    c = 0.11
    C5 = np.zeros((5, 5))
    C5[0, 0] = c
    i = np.arange(N)
    mode = np.cos(math.pi * k * (i + 0.5) / N)
    sh = [1, 1, 1]
    sh[axis] = N
    u = np.broadcast_to(mode.reshape(sh), shape).copy()
    out = op.stencil(u, C5, np.zeros(shape, np.uint8))
    lam = 1 - 4 * c * math.sin(math.pi * k / (2 * N)) ** 2
    np.testing.assert_allclose(out, lam * u, rtol=0, atol=1e-14)


def test_stencil_interface_mean_and_unconnected_wall():
    """Two half-spaces: flux across a connected interface uses the table
    entry (the arithmetic mean of the two sigmas); an unconnected pair of
    domains exchanges nothing."""
    shape = (6, 3, 3)
    ids = np.zeros(shape, np.uint8)
    ids[3:] = 1
    u = np.zeros(shape)
    u[:3] = 1.0
    C5 = np.zeros((5, 5))
    C5[0, 0], C5[1, 1] = 0.1, 0.3
    C5[0, 1] = C5[1, 0] = 0.2
    out = op.stencil(u, C5, ids)
    np.testing.assert_allclose(out[2], 1 - 0.2)
    np.testing.assert_allclose(out[3], 0.2)
    C5[0, 1] = C5[1, 0] = 0.0
    out = op.stencil(u, C5, ids)
    np.testing.assert_array_equal(out, u)


# ------------------------------------------------------------- reaction
def _reaction_only_model(ns=3):
    """One-domain model without diffusion: only buffer reactions act."""
    names = ["cyt"]
    sp = ["Ca", "B1", "B2"][:ns]
    bufs = [cs.Buffer(1, 0, 0.11, 0.11, 50.0), cs.Buffer(2, 0, 0.034, 0.238, 24.0)][:ns - 1]
    init = np.array([[2.0, 10.0, 3.0][:ns]])
    model = cs.SpeciesModel(domain_names=names, species_names=sp, diffusive=[[]], sigma=[[]],
                            init=init, buffers=[bufs])
    geom = cs.Geometry(voxels=np.zeros((4, 3, 10), np.uint8), h=6.0, domain_names=names,
                       domain_connections=np.zeros((1, 1), np.uint8))
    return geom, model


@need_ref_k
def test_reaction_matches_reaction_simd(refk):
    """A step with reactions only: the oracle equals the reference's
    reaction_SIMD run over every column and buffer, including its rule that
    a species not updated earlier in the step is overwritten (so free Ca
    keeps only the last buffer's term)."""
    geom, model = _reaction_only_model()
    dt = 1e-3
    orc = op.Oracle(geom, model, cs.Options(seed=1, dt=dt))
    rng = np.random.default_rng(4)
    shape = orc.n
    fields = [rng.uniform(0.5, 5.0, shape), rng.uniform(1, 40, shape), rng.uniform(1, 20, shape)]
    for s, f in enumerate(fields):
        orc.set_species(s, f)
    orc.step(1)
    nd, ns = 1, 3
    u1 = [np.ascontiguousarray(f.ravel()).copy() for f in fields]
    du = [x.copy() for x in u1]
    U1 = (dp * ns)(*[_ptr(a) for a in u1])
    DU = (dp * ns)(*[_ptr(a) for a in du])
    konT = np.zeros(ns * nd); koffT = np.zeros(ns * nd); totT = np.zeros(ns * nd)
    b0T = np.zeros(ns * nd, np.uint8); b1T = np.zeros(ns * nd, np.uint8)
    for bi, b in enumerate(model.buffers[0]):
        konT[bi], koffT[bi], totT[bi], b0T[bi], b1T[bi] = b.k_on, b.k_off, b.tot, b.bound, b.free
    ids = np.zeros(np.prod(shape) + 64, np.uint8)
    upd = np.zeros(ns, np.uint8)
    nx, ny, nz = shape
    refk.reaction_SIMD.argtypes = [C.POINTER(dp), C.POINTER(dp), dp, dp, dp, C.POINTER(C.c_uint8),
                                   C.POINTER(C.c_uint8), C.c_int, C.c_int, C.c_double,
                                   C.POINTER(C.c_uint8), C.c_int, C.c_int, C.POINTER(C.c_uint8)]
    for x in range(nx):
        for y in range(ny):
            for bi in range(2):
                refk.reaction_SIMD(U1, DU, _ptr(konT), _ptr(koffT), _ptr(totT), _ptr(b0T, C.c_uint8),
                                   _ptr(b1T, C.c_uint8), bi, (x * ny + y) * nz, dt, _ptr(ids, C.c_uint8),
                                   nd, nz, _ptr(upd, C.c_uint8))
    for s in range(ns):
        np.testing.assert_allclose(orc.species(s).ravel(), du[s], rtol=1e-15, atol=0)


def test_reaction_equilibrium_and_conservation():
    geom, model = _reaction_only_model()
    b1, b2 = model.buffers[0]
    ca = 0.7
    eq = lambda b: b.tot * ca / (ca + b.k_off / b.k_on)
    model.init = np.array([[ca, eq(b1), eq(b2)]])
    orc = op.Oracle(geom, model, cs.Options(seed=1, dt=1e-3))
    orc.step(3)
    for s, v in enumerate([ca, eq(b1), eq(b2)]):
        np.testing.assert_allclose(orc.species(s), v, rtol=1e-13)


# ------------------------------------------------------------- gating
class RefStates(C.Structure):
    _fields_ = [("N", C.c_uint), ("s0", C.POINTER(C.c_uint8))]


class RefRyR(C.Structure):
    """RyR_model_t of boundaryfluxes.h:31-56 in the DOUBLE build."""
    _fields_ = [("states", RefStates), ("Kd_open", C.c_double), ("k_min_open", C.c_double),
                ("k_max_open", C.c_double), ("n_open", C.c_double), ("Kd_close", C.c_double),
                ("k_min_close", C.c_double), ("k_max_close", C.c_double), ("n_close", C.c_double),
                ("i_unitary", C.c_double), ("t_close_ryrs", C.c_double),
                ("boundary_name", C.c_char * 20), ("species_name", C.c_char * 20),
                ("params", C.c_double * 10)]


def _gating_dyad():
    d = dyad.small_dyad(seed=21, num_ryr=9, num_pryr=3)
    import dataclasses
    fast = dict(Kd_open=0.02, k_max_open=80.0, k_min_close=60.0, k_max_close=60.0)
    model = dataclasses.replace(d.model, ryr=dataclasses.replace(d.model.ryr, **fast),
                                p_ryr=dataclasses.replace(d.model.p_ryr, n_close=-0.5, Kd_close=0.3))
    return d.geometry, model


@need_ref_f
def test_gating_matches_reference_functions(reff):
    """Stochastic init and per-step gating of the oracle (MT draw order)
    against the reference's init_RyR_model_states_stochastically and
    evaluate_RyR_stochastically fed the same Ca values and the same
    mtwist stream."""
    geom, model = _gating_dyad()
    seed = 99
    opts = cs.Options(seed=seed, rng="mt", dt_update_stoch=2.6e-5 * 3)
    orc = op.Oracle(geom, model, opts)
    setup = cs.Setup(geom, model, opts)
    P = setup.problem
    reff.mt_seed32new(seed)
    refs = []
    for f in range(2):
        F = P.fluxes[f]
        n = F.num_pairs
        st = np.zeros(n, np.uint8)
        r = RefRyR()
        r.states.N = n
        r.states.s0 = _ptr(st, C.c_uint8)
        for i in range(10):
            r.params[i] = F.params[i]
        refs.append((r, st, F))
    # ryr first, then p_ryr, both from the ryr boundary's to-voxel Ca
    ca0 = orc.species(0).ravel()
    ryr_to = np.array([P.fluxes[0].pairs[2 * i + 1] for i in range(P.fluxes[0].num_pairs)])
    vals = np.zeros(2 * len(ryr_to))
    vals[1::2] = ca0[ryr_to]
    reff.init_RyR_model_states_stochastically.argtypes = [C.POINTER(RefRyR), dp]
    for r, st, F in refs:
        reff.init_RyR_model_states_stochastically(C.byref(r), _ptr(vals))
    for k in range(2):
        np.testing.assert_array_equal(orc.states(k), refs[k][1])
        np.testing.assert_array_equal(np.array([P.fluxes[k].init_states[i] for i in range(P.fluxes[k].num_pairs)]),
                                      refs[k][1])
    reff.evaluate_RyR_stochastically.argtypes = [dp, C.POINTER(RefStates), C.c_double, C.c_double, dp]
    D = orc.info.stochastic_substep
    dt = orc.dt
    events = 0
    for ti in range(1, 121):
        ca = orc.species(0).ravel()
        if ti % D == 0:
            for r, st, F in refs:
                to = np.array([F.pairs[2 * i + 1] for i in range(F.num_pairs)])
                v = np.zeros(2 * len(to))
                v[1::2] = ca[to]
                before = st.copy()
                reff.evaluate_RyR_stochastically(r.params, C.byref(r.states), ti * dt, dt * D, _ptr(v))
                events += int(np.sum(before != st))
        orc.step(1)
        for k in range(2):
            np.testing.assert_array_equal(orc.states(k), refs[k][1], err_msg=f"time_ind {ti}")
    assert events > 0


# ------------------------------------------------------------- builder
def test_builder_matches_oracle_setup():
    for d, h in [(dyad.lattice_dyad(), 6.0), (dyad.small_dyad(n=(10, 9, 7), h=12.0), 6.0),
                 (dyad.small_dyad(n=(8, 8, 6), h=12.0), 3.0)]:
        opts = cs.Options(h=h, seed=5, dt_update_react=1e-4, dt_update_stoch=6e-5)
        s = cs.Setup(d.geometry, d.model, opts)
        o = op.Oracle(d.geometry, d.model, opts)
        P = s.problem
        assert s.n == o.n
        assert P.dt == o.dt
        assert list(P.substeps) == list(o.info.substeps)
        np.testing.assert_array_equal(np.array(P.coef), np.array(o.info.coef))
        assert P.num_buffers == o.info.num_all_buffers
        assert P.reaction_substep == o.info.reaction_substep
        assert P.stochastic_substep == o.info.stochastic_substep
        for k in range(2):
            np.testing.assert_array_equal(
                np.array([P.fluxes[k].init_states[i] for i in range(P.fluxes[k].num_pairs)]), o.states(k))


def test_dt_and_step_count_closed_forms():
    """lattice dyad at 6 nm: dt = 0.95 h^2 / (6 sigma_max) = 0.95*36/(6*220000) ms; the
    reference loop (t accumulated in double, calcium_sparks.c:190) runs
    ceil(10 ms / dt) = 385965 steps, counted here by an IEEE-double loop."""
    d = dyad.lattice_dyad()
    s = cs.Setup(d.geometry, d.model, cs.Options(seed=1))
    assert s.dt == pytest.approx(0.95 * 36 / (6 * 220000), rel=1e-15)
    t, n = 0.0, 0
    while t < 10.0:
        t += s.dt
        n += 1
    assert n == math.ceil(10.0 / s.dt) == 385965
    assert cs.host_lib().cs_count_steps(10.0, s.dt) == n
    # substeps k_s = floor(sigma_max / sigma_s) (species.c:281-294)
    assert list(s.problem.substeps)[:5] == [1, 5, 8, 1, 0]


def test_refined_discrete_pairs_use_centre_subvoxel():
    """A 12 nm channel pair jSR (x,y,z+1) -> cleft (x,y,z) run at 6 nm maps
    to the single fine pair (2x+1, 2y+1, 2z+2) -> (2x+1, 2y+1, 2z+1)
    (species.c:987-1020)."""
    d = dyad.small_dyad(n=(10, 9, 7), h=12.0, seed=11)
    s = cs.Setup(d.geometry, d.model, cs.Options(h=6.0, seed=1))
    n = s.n
    raw = d.geometry.boundary("ryr").pairs
    F = s.problem.fluxes[0]
    for i, (x0, y0, z0, x1, y1, z1) in enumerate(raw):
        f, t = F.pairs[2 * i], F.pairs[2 * i + 1]
        assert np.unravel_index(f, n) == (2 * x0 + 1, 2 * y0 + 1, 2 * z0)
        assert np.unravel_index(t, n) == (2 * x1 + 1, 2 * y1 + 1, 2 * z1 + 1)


def test_refined_density_pairs_expand_to_face():
    """A density pair at 12 nm run at 4 nm (3 subdivisions) becomes the 9
    fine pairs across the shared face."""
    import helpers
    d = helpers.with_serca(dyad.small_dyad(n=(8, 8, 6), h=12.0, seed=2))
    s = cs.Setup(d.geometry, d.model, cs.Options(h=4.0, seed=1))
    F = [s.problem.fluxes[f] for f in range(s.problem.num_fluxes) if s.problem.fluxes[f].kind == cs.FLUX_SERCA][0]
    raw = d.geometry.boundary("serca").pairs
    assert F.num_pairs == 9 * len(raw)
    x0, y0, z0, x1, y1, z1 = raw[0]
    got = {(np.unravel_index(F.pairs[2 * k], s.n), np.unravel_index(F.pairs[2 * k + 1], s.n)) for k in range(9)}
    want = {((3 * x0 + a, 3 * y0 + b, 3 * z0 + 2), (3 * x1 + a, 3 * y1 + b, 3 * z1)) for a in range(3) for b in range(3)}
    assert got == want


# ------------------------------------------------------------- linescan
# species.c:1960-2035 with gauss (utils.c:209-213) =
# exp(-x^2 / (2 sigma)) / sqrt(2 pi sigma), the normal density of variance
# sigma; scipy's norm.pdf is the independent check.

def _normal(x, var):
    from scipy.stats import norm
    return norm.pdf(x, scale=math.sqrt(var))


def _linescan_oracle(seed=5):
    d = dyad.small_dyad(seed=seed)
    orc = op.Oracle(d.geometry, d.model, cs.Options(seed=1))
    return d, orc


CONV = (2900.0, 3700.0, 9100.0)


@pytest.mark.parametrize("axis,offsets", [(0, (0, 2, -3)), (1, (1, 0, 2)), (2, (-2, 3, 0))])
def test_linescan_of_a_point_is_the_product_of_normal_densities(axis, offsets):
    d, orc = _linescan_oracle()
    n, h = orc.n, orc.info.h
    f = np.zeros(n)
    i, j, k = 7, 5, 9
    f[i, j, k] = 2.5
    orc.set_species(1, f)
    got = orc.linescan(1, axis, offsets, CONV)
    o = [(n[ax] // 2 + offsets[ax]) * h for ax in range(3)]
    c = (i * h, j * h, k * h)
    want = []
    for p in range(n[axis]):
        oo = list(o)
        oo[axis] = p * h
        want.append(2.5 * np.prod([_normal(oo[ax] - c[ax], CONV[ax]) for ax in range(3)]))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    orc.close()


def test_linescan_of_a_constant_field_factorises():
    d, orc = _linescan_oracle()
    n, h = orc.n, orc.info.h
    orc.set_species(2, np.full(n, 3.0))
    offsets = (1, -1, 2)
    got = orc.linescan(2, 1, offsets, CONV)
    sx = _normal((n[0] // 2 + offsets[0]) * h - np.arange(n[0]) * h, CONV[0]).sum()
    sz = _normal((n[2] // 2 + offsets[2]) * h - np.arange(n[2]) * h, CONV[2]).sum()
    want = [3.0 * sx * sz * _normal(p * h - np.arange(n[1]) * h, CONV[1]).sum() for p in range(n[1])]
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    orc.close()


def test_linescan_domain_rule_compares_ids_with_the_count():
    """With k listed domains the reference keeps the voxels whose id is
    below k (species.c:2015-2024), whatever the listed names are."""
    d, orc = _linescan_oracle()
    ids = np.asarray(d.geometry.voxels)
    assert ids.shape == tuple(orc.n)
    rng = np.random.default_rng(3)
    f = rng.random(orc.n)
    orc.set_species(0, f)
    full = orc.linescan(0, 2, (0, 0, 0), CONV)
    for k in (1, 2, 3):
        orc.set_species(0, np.where(ids < k, f, 0.0))
        want = orc.linescan(0, 2, (0, 0, 0), CONV)
        orc.set_species(0, f)
        np.testing.assert_allclose(orc.linescan(0, 2, (0, 0, 0), CONV, num_domains=k), want, rtol=1e-13)
    orc.set_species(0, np.where(ids >= 1, f, 0.0))
    assert np.all(orc.linescan(0, 0, (0, 0, 0), CONV, num_domains=1) == 0.0)
    assert np.any(full > 0)
    orc.close()


def test_linescan_separable_evaluation_matches_direct_sum():
    """Projection with the transverse weights, then a 1-D convolution --
    the order the device path uses -- equals the direct triple sum."""
    d, orc = _linescan_oracle(seed=6)
    n, h = orc.n, orc.info.h
    f = np.random.default_rng(4).random(n)
    orc.set_species(3, f)
    for axis, offsets in [(0, (0, 1, 1)), (1, (2, 0, -1)), (2, (0, 0, 0))]:
        o = [(n[ax] // 2 + offsets[ax]) * h for ax in range(3)]
        w = [_normal(o[ax] - np.arange(n[ax]) * h, CONV[ax]) for ax in range(3)]
        others = [ax for ax in range(3) if ax != axis]
        proj = np.moveaxis(f, axis, 0)
        proj = np.tensordot(proj, np.multiply.outer(w[others[0]], w[others[1]]), axes=([1, 2], [0, 1]))
        k = np.arange(n[axis]) * h
        want = [float(np.dot(proj, _normal(p - k, CONV[axis]))) for p in k]
        np.testing.assert_allclose(orc.linescan(3, axis, offsets, CONV), want, rtol=1e-12)
    orc.close()
