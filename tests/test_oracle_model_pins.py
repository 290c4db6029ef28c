"""Pins of the oracle's whole-model functions (CPU only).

Each test fixes an oracle function by a property the reference's
definition implies, chosen so that a plausible slip fails it:

* coefficient table (species.c:1547-1560): diffusion of a Neumann
  eigenmode through the full oracle step decays by 1 - 4 c sin^2 exactly
  once every k_s steps with c = dt k_s sigma / h^2 (a missing k_s, a wrong
  power of h or a substep that fires every step fails it); an interface
  between two connected domains with different sigma passes the arithmetic
  mean (hand-computed literals); a species listed only in a prefix of a
  domain's diffusive list reads sigma[d][dsi] by its global diffusive
  index;
* output reductions (species.c:2806-2863): constant and linear fields give
  exact per-domain min, max and sum*dV; max starts from FLT_MIN, so an
  all-negative domain reports FLT_MIN, and an empty domain reports the
  start values;
* Dirichlet faces (-f, species.c:1046-1064): after a step the face values
  of species 0 are init[domain][0] and nothing else differs from the same
  step without -f;
* fixed domain species (species.c:1066-1077): the value stays
  init[domain][species] in that domain only;
* the coupled step order (calcium_sparks.c:196-227): with RyR and SERCA
  fluxes between inner voxels, buffers and unequal substeps, total
  calcium over every slot is conserved to round-off.
"""
import math

import numpy as np
import pytest

import csparks as cs
import dyad
import oracle_py as op
from helpers import with_serca

FLT_MIN = np.float32(np.finfo(np.float32).tiny).item()
FLT_MAX = np.float32(np.finfo(np.float32).max).item()


def _one_domain_model(sigmas, init=None):
    """One domain, len(sigmas) species, each diffusive with its sigma, no
    buffers or channels."""
    ns = len(sigmas)
    init = np.ones((1, ns)) if init is None else np.asarray(init, dtype=float).reshape(1, ns)
    return cs.SpeciesModel(domain_names=["cyt"], species_names=[f"s{i}" for i in range(ns)],
                           diffusive=[list(range(ns))], sigma=[list(map(float, sigmas))], init=init,
                           buffers=[[]])


def _box(shape, ids=None, names=("cyt",), conn=None):
    vox = np.zeros(shape, dtype=np.uint8) if ids is None else ids.astype(np.uint8)
    nd = len(names)
    conn = np.ones((nd, nd), dtype=np.uint8) if conn is None else conn
    return cs.Geometry(voxels=vox, h=6.0, domain_names=list(names), domain_connections=conn, boundaries=[])


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_eigenmode_decays_once_per_substep_with_k_scaled_coefficient(axis):
    """sigma_max = 200000 sets dt; species 1 with sigma 30000 gets
    k_s = floor(0.95 h^2/(6*30000) / dt) = floor(20/3) = 6.  A cosine mode
    of species 1 stays unchanged for five steps and decays at step 6 by
    1 - 4 c sin^2(pi k/(2N)), c = dt * 6 * 30000 / 36."""
    shape = [6, 7, 8]
    N = shape[axis]
    g = _box(tuple(shape))
    m = _one_domain_model([200000.0, 30000.0])
    o = op.Oracle(g, m, cs.Options(h=6.0, seed=1))
    dt = 0.95 * 36.0 / (6 * 200000.0)
    assert o.dt == pytest.approx(dt, rel=1e-15)
    assert list(o.info.substeps)[:2] == [1, 6]
    i = np.arange(N)
    mode = np.cos(np.pi * 1 * (i + 0.5) / N)
    shp = [1, 1, 1]
    shp[axis] = N
    u = 1.0 + 0.25 * np.broadcast_to(mode.reshape(shp), tuple(shape)).copy()
    o.set_species(1, u)
    for step in range(1, 6):
        o.step(1)
        np.testing.assert_array_equal(o.species(1), u, err_msg=f"species 1 changed at step {step}")
    o.step(1)
    c = dt * 6 * 30000.0 / 36.0
    factor = 1.0 - 4.0 * c * math.sin(math.pi / (2 * N)) ** 2
    want = 1.0 + 0.25 * factor * np.broadcast_to(mode.reshape(shp), tuple(shape))
    np.testing.assert_allclose(o.species(1), want, rtol=0, atol=1e-15)
    o.close()


def test_interface_coefficient_is_the_arithmetic_mean():
    """Two connected domains along x with sigma 100000 | 300000 (one
    species, k = 1): a unit excess in the last voxel of domain 0 moves
    dt (100000 + 300000)/2 / h^2 into the first voxel of domain 1 in one
    step; with dt = 0.95*36/(6*300000) that is 0.95*36/(6*300000)*200000/36
    = 0.95/9 = 0.105555... (hand computation)."""
    ids = np.zeros((4, 1, 1), dtype=np.uint8)
    ids[2:] = 1
    g = _box((4, 1, 1), ids=ids, names=("a", "b"))
    m = cs.SpeciesModel(domain_names=["a", "b"], species_names=["s0"], diffusive=[[0], [0]],
                        sigma=[[100000.0], [300000.0]], init=np.zeros((2, 1)), buffers=[[], []])
    o = op.Oracle(g, m, cs.Options(h=6.0, seed=1))
    u = np.zeros((4, 1, 1))
    u[1] = 1.0
    o.set_species(0, u)
    o.step(1)
    got = o.species(0).ravel()
    assert got[2] == pytest.approx(0.95 / 9.0, rel=1e-14)
    # inside domain 0 the coefficient is dt*100000/36 = 0.95/18
    assert got[0] == pytest.approx(0.95 / 18.0, rel=1e-14)
    assert got[3] == 0.0
    assert got.sum() == pytest.approx(1.0, rel=1e-15)
    # unconnected, nothing crosses
    g2 = _box((4, 1, 1), ids=ids, names=("a", "b"), conn=np.eye(2, dtype=np.uint8))
    o2 = op.Oracle(g2, m, cs.Options(h=6.0, seed=1))
    o2.set_species(0, u)
    o2.step(1)
    assert o2.species(0).ravel()[2] == 0.0
    o.close()
    o2.close()


def test_sigma_is_read_by_global_diffusive_index():
    """species.c:1547-1560 reads sigma[d][dsi] with dsi the position of the
    species among all diffusive species.  Domain a lists species (0, 1),
    domain b lists only species 0: for species 1 (dsi 1) domain b has no
    entry, so C[b][b] = 0 and C[a][b] = 0 although the domains connect,
    while C[a][a] = dt sigma_a1 / h^2."""
    ids = np.zeros((4, 1, 1), dtype=np.uint8)
    ids[2:] = 1
    g = _box((4, 1, 1), ids=ids, names=("a", "b"))
    m = cs.SpeciesModel(domain_names=["a", "b"], species_names=["s0", "s1"], diffusive=[[0, 1], [0]],
                        sigma=[[200000.0, 100000.0], [200000.0]], init=np.zeros((2, 2)), buffers=[[], []])
    o = op.Oracle(g, m, cs.Options(h=6.0, seed=1))
    C = np.array(o.info.coef[1]).reshape(5, 5)
    dt = 0.95 * 36 / (6 * 200000.0)
    # k_1 = floor(dt_1/dt), dt_1 = 0.95*36/(6*100000) = 2 dt
    k1 = o.info.substeps[1]
    assert k1 == 2
    assert C[0, 0] == pytest.approx(dt * k1 * 100000.0 / 36.0, rel=1e-15)
    assert C[1, 1] == 0.0 and C[0, 1] == 0.0 and C[1, 0] == 0.0
    o.close()


def _stats_oracle():
    """6x5x4 box, domain 0 for x < 3 and domain 1 for x >= 3, domain 2
    listed but empty; two species."""
    ids = np.zeros((6, 5, 4), dtype=np.uint8)
    ids[3:] = 1
    g = _box((6, 5, 4), ids=ids, names=("a", "b", "c"), conn=np.zeros((3, 3), dtype=np.uint8))
    m = cs.SpeciesModel(domain_names=["a", "b", "c"], species_names=["s0", "s1"], diffusive=[[], [], []],
                        sigma=[[], [], []], init=np.zeros((3, 2)), buffers=[[], [], []])
    return op.Oracle(g, m, cs.Options(h=6.0, seed=1)), ids


def test_stats_of_constant_and_linear_fields():
    o, ids = _stats_oracle()
    dV = 216.0
    x = np.arange(6).reshape(6, 1, 1) * np.ones((6, 5, 4))
    o.set_species(0, np.full((6, 5, 4), 2.5))
    o.set_species(1, 3.0 + 2.0 * x)          # x = 0..5
    mn, mx, tot = o.stats()
    # species 0, constant 2.5: 60 voxels per domain
    assert mn[0, 0] == mx[0, 0] == 2.5 and mn[1, 0] == mx[1, 0] == 2.5
    assert tot[0, 0] == 2.5 * 60 * dV and tot[1, 0] == 2.5 * 60 * dV
    # species 1, 3 + 2x: domain a has x = 0,1,2 (20 voxels each), b x = 3,4,5
    assert (mn[0, 1], mx[0, 1]) == (3.0, 7.0) and (mn[1, 1], mx[1, 1]) == (9.0, 13.0)
    assert tot[0, 1] == (3 + 5 + 7) * 20 * dV and tot[1, 1] == (9 + 11 + 13) * 20 * dV
    # empty domain c keeps the start values
    assert mn[2, 0] == FLT_MAX and mx[2, 0] == FLT_MIN and tot[2, 0] == 0.0
    o.close()


def test_stats_max_starts_from_flt_min():
    """An all-negative domain reports max = FLT_MIN (the smallest positive
    float), not its largest value (species.c:2808-2814)."""
    o, ids = _stats_oracle()
    v = np.where(ids == 0, -1.0, 4.0).astype(float)
    o.set_species(0, v)
    o.set_species(1, np.zeros((6, 5, 4)))
    mn, mx, tot = o.stats()
    assert mn[0, 0] == -1.0 and mx[0, 0] == FLT_MIN
    assert mn[1, 0] == 4.0 and mx[1, 0] == 4.0
    # zero fields: max = FLT_MIN as well, min = 0
    assert mx[0, 1] == FLT_MIN and mn[0, 1] == 0.0
    o.close()


def _perturbed(o, rng):
    for s in range(o.ns):
        u = o.species(s)
        o.set_species(s, u * (1.0 + 0.05 * rng.random(u.shape)))


def test_dirichlet_faces_reset_species0_only():
    d = dyad.small_dyad(seed=2)
    base = dict(seed=3, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0})
    a = op.Oracle(d.geometry, d.model, cs.Options(force_dirichlet=True, **base))
    b = op.Oracle(d.geometry, d.model, cs.Options(**base))
    rng = np.random.default_rng(1)
    _perturbed(a, rng)
    for s in range(a.ns):
        b.set_species(s, a.species(s))
    a.step(1)
    b.step(1)
    ids = d.geometry.voxels
    face = np.zeros(ids.shape, dtype=bool)
    face[0], face[-1], face[:, 0], face[:, -1], face[:, :, 0], face[:, :, -1] = (True,) * 6
    init0 = np.asarray(d.model.init)[ids, 0]
    ua, ub = a.species(0), b.species(0)
    np.testing.assert_array_equal(ua[face], init0[face])
    assert np.any(ub[face] != init0[face])
    np.testing.assert_array_equal(ua[~face], ub[~face])
    for s in range(1, a.ns):
        np.testing.assert_array_equal(a.species(s), b.species(s))
    a.close()
    b.close()


def test_fixed_domain_species_stay_at_init():
    """(cleft, Fluo) fixed: Fluo in the cleft stays at its resting value
    while an open RyR drives Fluo binding elsewhere; Fluo in the cytosol
    moves."""
    import dataclasses
    d = dyad.small_dyad(seed=5)
    cleft = d.model.domain_names.index("cleft")
    model = dataclasses.replace(d.model, fixed=[(cleft, 1)])
    o = op.Oracle(d.geometry, model, cs.Options(seed=2, open={"ryr": [d.open_index]}, t_close={"ryr": 1.0}))
    ids = d.geometry.voxels
    f0 = o.species(1).copy()
    o.step(40)
    f = o.species(1)
    rest = np.asarray(model.init)[cleft, 1]
    np.testing.assert_array_equal(f[ids == cleft], rest)
    cyt = ids == d.model.domain_names.index("cyt")
    assert np.max(np.abs(f[cyt] - f0[cyt])) > 1e-9
    o.close()


def test_coupled_step_conserves_total_calcium():
    """RyR (inner, jSR -> cleft) and SERCA (inner, cytosol -> nSR) fluxes,
    four buffers with reactions every step, Fluo/CMDN diffusing every 5th
    and 8th step, no Dirichlet: the sum of every slot (free plus bound
    calcium) times dV is invariant."""
    d = with_serca(dyad.small_dyad(n=(16, 14, 12), seed=7))
    o = op.Oracle(d.geometry, d.model, cs.Options(seed=9, open={"ryr": [0, 1, 2]}))
    t0 = o.stats()[2]
    o.step(123)
    t1 = o.stats()[2]
    assert abs(t1.sum() - t0.sum()) <= 1e-13 * t0.sum()
    # a nontrivial run: calcium left the jSR and entered the nSR
    jsr, nsr = (d.model.domain_names.index(n) for n in ("jsr", "nsr"))
    assert t0[jsr].sum() - t1[jsr].sum() > 1e-6 * t0[jsr].sum()
    assert t1[nsr].sum() > t0[nsr].sum()
    o.close()


def test_interleaved_oracle_runs_keep_their_own_generators():
    """Each oracle run draws from its own Mersenne Twister: two runs stepped
    alternately give the channel trajectories each gives alone."""
    import dataclasses
    d = dyad.small_dyad(seed=4)
    kin = dict(Kd_open=2.0, k_max_open=300.0, k_min_close=300.0, k_max_close=300.0)
    m = dataclasses.replace(d.model, ryr=dataclasses.replace(d.model.ryr, **kin),
                            p_ryr=dataclasses.replace(d.model.p_ryr, **kin))
    o1 = cs.Options(seed=11, rng="mt", open={"ryr": [d.open_index]})
    o2 = cs.Options(seed=12, rng="mt", open={"ryr": [d.open_index]})

    def alone(o):
        r = op.Oracle(d.geometry, m, o)
        out = []
        for _ in range(30):
            r.step(20)
            out.append(np.concatenate([r.states(0), r.states(1)]))
        r.close()
        return out

    a1, a2 = alone(o1), alone(o2)
    r1, r2 = op.Oracle(d.geometry, m, o1), op.Oracle(d.geometry, m, o2)
    flips = 0
    for k in range(30):
        r1.step(20)
        r2.step(20)
        np.testing.assert_array_equal(np.concatenate([r1.states(0), r1.states(1)]), a1[k])
        np.testing.assert_array_equal(np.concatenate([r2.states(0), r2.states(1)]), a2[k])
        flips += int(np.sum(a1[k] != a1[k - 1])) if k else 0
    assert flips > 0
    r1.close()
    r2.close()
