"""Synthetic dyad generator in the reference's input formats.

The reference ships a handful of 12 nm geometries and no generator (its
README's converter is absent), so the benchmark workloads are built here:
a T-tubule slab at the bottom, a cleft above it under the junctional SR
footprint, a jSR slab above the cleft, cytosol elsewhere; RyR and pRyR
channels are discrete boundaries from the first jSR layer into the last
cleft layer.  Domain order and species slots follow the shipped species
files: domains cyt, cleft, jsr, nsr, tt; species Ca, Fluo, CMDN-CSQN,
ATP, TRPN (SR Ca shares slot 0, CSQN shares slot 2 with CMDN).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from csparks import (Boundary, Buffer, ChannelModel, Geometry, SpeciesModel)

CYT, CLEFT, JSR, NSR, TT = 0, 1, 2, 3, 4
DOMAINS = ["cyt", "cleft", "jsr", "nsr", "tt"]
SPECIES = ["Ca", "Fluo", "CMDN-CSQN", "ATP", "TRPN"]


def hill_as_power_law(k_max: float, K: float, n: float, k_min: float = 1e-6) -> Tuple[float, float, float, float]:
    """Express k = k_max c^n/(c^n+K^n) in the reference's clamped power law
    min(max((c/Kd)^n, k_min), k_max): Kd = K k_max^(-1/n) matches both
    asymptotes.  Returns (Kd_open, n_open, k_max_open, k_min_open)."""
    return K * k_max ** (-1.0 / n), n, k_max, k_min


def equilibrium(tot: float, kon: float, koff: float, ca: float) -> float:
    return tot * ca / (ca + koff / kon)


def dyad_species_model(ca_cyt: float = 0.1, ca_sr: float = 1300.0, K_nonp: float = 70.0,
                       K_p: float = 35.0, k_open_max: float = 0.3, n_open: float = 2.8,
                       k_close: float = 0.5, i_pA_at_1mM: float = 0.15,
                       with_pryr: bool = True) -> SpeciesModel:
    """Species file of the benchmark dyads: 7 species in 5 slots.

    Cytosol/cleft: Ca (D=220 um^2/s), Fluo-4 (D=42, 50 uM, kon 0.11, koff
    0.11), CMDN (D=25, 24 uM, kon 0.034, koff 0.238), ATP (D=140, 5 mM,
    kon 0.225, koff 45), TRPN (immobile, 70 uM, kon 0.0327, koff 0.0196).
    SR: Ca (D=60, 1300 uM), CSQN (immobile, 30 mM, kon 0.1, koff 63).
    Buffers start at equilibrium.  RyR gating: Hill opening with K=70 uM
    (pRyR 35 uM), k_max 0.3/ms, n=2.8, closing 0.5/ms; 0.15 pA at 1 mM.
    """
    ns, nd = len(SPECIES), len(DOMAINS)
    cyt_bufs = [Buffer(1, 0, 0.11, 0.11, 50.0), Buffer(2, 0, 0.034, 0.238, 24.0),
                Buffer(3, 0, 0.225, 45.0, 5000.0), Buffer(4, 0, 0.0327, 0.0196, 70.0)]
    sr_bufs = [Buffer(2, 0, 0.1, 63.0, 30000.0)]
    init = np.zeros((nd, ns))
    for d in (CYT, CLEFT):
        init[d, 0] = ca_cyt
        for b in cyt_bufs:
            init[d, b.bound] = equilibrium(b.tot, b.k_on, b.k_off, ca_cyt)
    for d in (JSR, NSR):
        init[d, 0] = ca_sr
        for b in sr_bufs:
            init[d, b.bound] = equilibrium(b.tot, b.k_on, b.k_off, ca_sr)
    cyt_diff, cyt_sigma = [0, 1, 2, 3], [220000.0, 42000.0, 25000.0, 140000.0]
    diffusive = [cyt_diff, cyt_diff, [0], [0], []]
    sigma = [cyt_sigma, cyt_sigma, [60000.0], [60000.0], []]
    buffers = [cyt_bufs, cyt_bufs, sr_bufs, sr_bufs, []]
    # the reference normalises i_unitary to a 1299.9 uM gradient
    i_unitary = i_pA_at_1mM * (1300.0 - 0.1) / 1000.0

    def chan(K, boundary):
        Kd, n, kmax, kmin = hill_as_power_law(k_open_max, K, n_open)
        return ChannelModel(Kd_open=Kd, k_min_open=kmin, k_max_open=kmax, n_open=n,
                            Kd_close=1.0, k_min_close=k_close, k_max_close=k_close, n_close=0.0,
                            i_unitary=i_unitary, boundary=boundary, species="Ca")

    return SpeciesModel(domain_names=list(DOMAINS), species_names=list(SPECIES),
                        diffusive=diffusive, sigma=sigma, init=init, buffers=buffers,
                        ryr=chan(K_nonp, "ryr"), p_ryr=chan(K_p, "p_ryr") if with_pryr else None,
                        serca=None)


def connections() -> np.ndarray:
    c = np.zeros((len(DOMAINS), len(DOMAINS)), dtype=np.uint8)
    for a, b in ((CYT, CLEFT), (JSR, NSR)):
        c[a, b] = c[b, a] = 1
    return c


@dataclass
class DyadLayout:
    n: Tuple[int, int, int]        # voxels
    h: float                       # nm
    tt: int                        # T-tubule slab thickness (voxels)
    cleft: int                     # cleft thickness (voxels)
    jsr: int                       # jSR slab thickness (voxels)
    footprint: Tuple[int, int, int, int]  # x0, x1, y0, y1 of the jSR/cleft


def dyad_geometry(layout: DyadLayout, ryr_xy: Sequence[Tuple[int, int]], pryr_mask: Sequence[bool],
                  always_pryr: bool = False) -> Geometry:
    """Voxel geometry + discrete 'ryr' / 'p_ryr' boundaries (jSR -> cleft);
    with always_pryr the 'p_ryr' boundary exists even without channels."""
    nx, ny, nz = layout.n
    vox = np.full((nx, ny, nz), CYT, dtype=np.uint8)
    vox[:, :, :layout.tt] = TT
    x0, x1, y0, y1 = layout.footprint
    zc0, zc1 = layout.tt, layout.tt + layout.cleft
    zj0, zj1 = zc1, zc1 + layout.jsr
    vox[x0:x1, y0:y1, zc0:zc1] = CLEFT
    vox[x0:x1, y0:y1, zj0:zj1] = JSR
    ryr, pryr = [], []
    for (x, y), p in zip(ryr_xy, pryr_mask):
        if not (x0 <= x < x1 and y0 <= y < y1):
            raise ValueError("channel outside the jSR footprint")
        (pryr if p else ryr).append((x, y, zj0, x, y, zc1 - 1))
    bnds = [Boundary("ryr", "discrete", np.array(ryr, dtype=np.uint16).reshape(-1, 6))]
    if pryr or always_pryr:
        bnds.append(Boundary("p_ryr", "discrete", np.array(pryr, dtype=np.uint16).reshape(-1, 6)))
    return Geometry(voxels=vox, h=layout.h, domain_names=list(DOMAINS),
                    domain_connections=connections(), boundaries=bnds)


def lattice(nside: int, pitch: int, center: Tuple[int, int]) -> List[Tuple[int, int]]:
    cx, cy = center
    off = (nside - 1) // 2
    return [(cx + (i - off) * pitch, cy + (j - off) * pitch) for i in range(nside) for j in range(nside)]


def poisson_disk(rng: np.random.Generator, count: int, radius_vox: float, min_dist_vox: float,
                 center: Tuple[int, int], max_tries: int = 200000) -> List[Tuple[int, int]]:
    """Dart throwing on the voxel grid inside a disk, rejecting points
    closer than min_dist_vox to an accepted one."""
    pts: List[Tuple[int, int]] = []
    cx, cy = center
    tries = 0
    while len(pts) < count and tries < max_tries:
        tries += 1
        r = radius_vox * math.sqrt(rng.random())
        a = 2 * math.pi * rng.random()
        x, y = int(round(cx + r * math.cos(a))), int(round(cy + r * math.sin(a)))
        if all((x - px) ** 2 + (y - py) ** 2 >= min_dist_vox ** 2 for px, py in pts):
            pts.append((x, y))
    if len(pts) < count:
        raise RuntimeError("Poisson-disk sampling could not place all channels")
    return pts


@dataclass
class Dyad:
    geometry: Geometry
    model: SpeciesModel
    open_index: Optional[int]      # index of the forced-open RyR in boundary "ryr"
    description: str


def lattice_dyad(seed: int = 1, num_pryr: int = 10) -> Dyad:
    """384x384x150 nm at 6 nm (64x64x25): T-tubule 0-30 nm, cleft 30-42 nm
    and jSR 42-72 nm under the central 240x240 nm; 49 RyRs on a 7x7 lattice
    at 30 nm pitch, num_pryr of them phosphorylated (chosen by seed); the
    central RyR is the one forced open."""
    lay = DyadLayout(n=(64, 64, 25), h=6.0, tt=5, cleft=2, jsr=5, footprint=(12, 52, 12, 52))
    pts = lattice(7, 5, (32, 32))
    rng = np.random.default_rng(seed)
    centre = 24
    cand = [i for i in range(len(pts)) if i != centre]
    chosen = set(rng.choice(cand, size=num_pryr, replace=False).tolist()) if num_pryr else set()
    mask = [i in chosen for i in range(len(pts))]
    geom = dyad_geometry(lay, pts, mask)
    ryr_idx = [i for i in range(len(pts)) if not mask[i]]
    model = dyad_species_model(with_pryr=num_pryr > 0)
    return Dyad(geom, model, ryr_idx.index(centre), "lattice dyad 64x64x25 @6nm, 49 RyR")


def cluster_dyad(seed: int = 1, n: Tuple[int, int, int] = (400, 400, 200), h: float = 3.0,
            num_ryr: int = 100, pryr_fraction: float = 0.3) -> Dyad:
    """1200x1200x600 nm at 3 nm (400x400x200 = 32M voxels): cleft 12 nm
    (4 voxels), jSR slab 30 nm over a 600x600 nm footprint, 100 RyRs by
    seeded Poisson-disk sampling (>= 25 nm apart) inside a 400 nm disk,
    30% phosphorylated."""
    nx, ny, nz = n
    scale = nx / 400.0
    tt = max(1, int(round(30.0 / h)))
    cleft = max(1, int(round(12.0 / h)))
    jsr = max(1, int(round(30.0 / h)))
    fp = int(round(600.0 / h * scale))
    cx, cy = nx // 2, ny // 2
    lay = DyadLayout(n=n, h=h, tt=tt, cleft=cleft, jsr=jsr,
                     footprint=(cx - fp // 2, cx + fp // 2, cy - fp // 2, cy + fp // 2))
    rng = np.random.default_rng(seed)
    pts = poisson_disk(rng, num_ryr, radius_vox=200.0 / h * scale, min_dist_vox=25.0 / h * scale,
                       center=(cx, cy))
    npryr = int(round(pryr_fraction * num_ryr))
    chosen = set(rng.choice(range(1, num_ryr), size=npryr, replace=False).tolist())
    mask = [i in chosen for i in range(num_ryr)]
    geom = dyad_geometry(lay, pts, mask)
    model = dyad_species_model()
    return Dyad(geom, model, 0, f"cluster dyad {nx}x{ny}x{nz} @{h}nm, {num_ryr} RyR")


def small_dyad(n: Tuple[int, int, int] = (20, 18, 14), h: float = 6.0, seed: int = 3,
               num_ryr: int = 9, num_pryr: int = 3) -> Dyad:
    """A few-thousand-voxel dyad for parity tests that the oracle finishes
    in seconds; same structure as lattice_dyad."""
    nx, ny, nz = n
    mx, my = max(1, min(3, nx // 4)), max(1, min(3, ny // 4))
    lay = DyadLayout(n=n, h=h, tt=2, cleft=2, jsr=3, footprint=(mx, nx - mx, my, ny - my))
    rng = np.random.default_rng(seed)
    side = int(math.ceil(math.sqrt(num_ryr)))
    room = min(nx - 2 * mx, ny - 2 * my) - 1
    pitch = 2 if 2 * (side - 1) < room else 1
    pts = lattice(side, pitch, (nx // 2, ny // 2))[:num_ryr]
    chosen = set(rng.choice(range(1, num_ryr), size=num_pryr, replace=False).tolist()) if num_pryr else set()
    mask = [i in chosen for i in range(num_ryr)]
    geom = dyad_geometry(lay, pts, mask)
    model = dyad_species_model(with_pryr=num_pryr > 0)
    return Dyad(geom, model, 0, f"small {n} @{h}nm")


def fast_spark_dyad(seed: int = 3, num_ryr: int = 16, num_pryr: int = 4) -> Dyad:
    """small_dyad (20x18x14 at 6 nm) with RyR kinetics several hundred
    times faster than the shipped ones (open at up to 200/ms above ~5 uM
    cleft Ca, close at 50/ms), so a spark starts, recruits and ends within
    a few tenths of a millisecond: spark statistics over hundreds of seeds
    then stay affordable for the CPU oracle.  With the first RyR forced
    open, about half the realisations recruit the whole cluster and half
    end with the first channel."""
    import dataclasses
    d = small_dyad(seed=seed, num_ryr=num_ryr, num_pryr=num_pryr)
    fast = dict(Kd_open=5.0, k_max_open=200.0, k_min_close=50.0, k_max_close=50.0)
    m = d.model
    model = dataclasses.replace(m, ryr=dataclasses.replace(m.ryr, **fast),
                                p_ryr=dataclasses.replace(m.p_ryr, **fast) if m.p_ryr else None)
    return Dyad(d.geometry, model, d.open_index, f"fast-gating {d.description}")


def _cluster(rng: np.random.Generator, count: int, radius_vox: float, h: float,
             center: Tuple[int, int]) -> List[Tuple[int, int]]:
    """`count` distinct voxels in a disk by random sequential adsorption:
    candidates are drawn uniformly in the disk and kept when no kept point
    lies closer than the spacing.  The spacing starts at 25 nm, or lower
    when the disk cannot hold `count` points at that spacing (about 0.55
    of the close packing of disks), and shrinks by 10% whenever 40 draws
    per point have not placed them all."""
    cx, cy = center
    area = math.pi * radius_vox ** 2
    spacing = min(25.0 / h, 0.9 * math.sqrt(0.55 * area / (math.pi / 4) / count))
    cells = {(x, y) for x in range(int(cx - radius_vox) - 1, int(cx + radius_vox) + 2)
             for y in range(int(cy - radius_vox) - 1, int(cy + radius_vox) + 2)
             if (x - cx) ** 2 + (y - cy) ** 2 <= radius_vox ** 2}
    if len(cells) < count:
        raise RuntimeError("the cluster disk has fewer voxels than channels")
    while True:
        d2 = max(spacing, 1.0) ** 2
        pts: List[Tuple[int, int]] = []
        arr = np.empty((count, 2))
        r = radius_vox * np.sqrt(rng.random(40 * count))
        a = 2 * math.pi * rng.random(40 * count)
        xs = np.rint(cx + r * np.cos(a)).astype(int)
        ys = np.rint(cy + r * np.sin(a)).astype(int)
        for x, y in zip(xs, ys):
            if (x, y) not in cells:
                continue
            k = len(pts)
            if k and np.min((arr[:k, 0] - x) ** 2 + (arr[:k, 1] - y) ** 2) < d2:
                continue
            arr[k] = (x, y)
            pts.append((int(x), int(y)))
            if len(pts) == count:
                return pts
        spacing *= 0.9


def ensemble_dyad(seed: int, h: float = 6.0, edge_nm: Tuple[float, float] = (256.0, 512.0),
                  ryr: Tuple[int, int] = (9, 100), pryr_max: float = 0.5, depth_nm: float = 150.0) -> Dyad:
    """One realisation of the ensemble workload, drawn from `seed`: a square
    dyad of edge L ~ U[edge_nm] and depth depth_nm at voxel size h, with
    the lattice dyad's layers (T-tubule 30 nm, cleft 12 nm, jSR 30 nm); a
    cluster of N ~ U{ryr} RyRs by Poisson-disk sampling (_cluster) in a
    disk sized to about one RyR per (27 nm)^2; a fraction ~ U[0, pryr_max] of them phosphorylated (the p_ryr
    boundary exists even when empty, so every realisation has the same
    fluxes); the jSR/cleft footprint covers the cluster and at least 62.5%
    of the edge.  open_index is the non-phosphorylated RyR nearest the
    cluster centre, the one that starts the spark."""
    rng = np.random.default_rng(seed)
    L = float(rng.uniform(*edge_nm))
    nxy = max(8, int(round(L / h)))
    nz = max(6, int(round(depth_nm / h)))
    count = int(rng.integers(ryr[0], ryr[1] + 1))
    frac = float(rng.uniform(0.0, pryr_max))
    tt, cleft, jsr = max(1, int(round(30.0 / h))), max(1, int(round(12.0 / h))), max(1, int(round(30.0 / h)))
    if tt + cleft + jsr + 1 > nz:
        raise ValueError("dyad too shallow for its layers")
    c = nxy // 2
    radius_nm = min(27.0 * math.sqrt(count / math.pi) * 1.15, 0.5 * (nxy - 4) * h)
    pts = _cluster(rng, count, radius_nm / h, h, (c, c))
    half = max(int(math.ceil(radius_nm / h)) + 2, int(round(0.625 * nxy / 2)))
    half = min(half, c - 1, nxy - c - 1)
    lay = DyadLayout(n=(nxy, nxy, nz), h=h, tt=tt, cleft=cleft, jsr=jsr,
                     footprint=(c - half, c + half, c - half, c + half))
    npryr = int(round(frac * count))
    dist = [(x - c) ** 2 + (y - c) ** 2 for x, y in pts]
    trigger = int(np.argmin(dist))
    cand = [i for i in range(count) if i != trigger]
    chosen = set(rng.choice(cand, size=npryr, replace=False).tolist()) if npryr else set()
    mask = [i in chosen for i in range(count)]
    geom = dyad_geometry(lay, pts, mask, always_pryr=True)
    ryr_idx = [i for i in range(count) if not mask[i]]
    model = dyad_species_model(with_pryr=True)
    return Dyad(geom, model, ryr_idx.index(trigger),
                f"ensemble dyad {nxy}x{nxy}x{nz} @{h}nm, {count} RyR ({npryr} pRyR)")
