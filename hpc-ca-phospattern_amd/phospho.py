"""Phosphorylation-pattern converter: a 2-D RyR mask -> a geometry with
'ryr' and 'p_ryr' discrete boundaries.

The reference ships each dyad's RyR layout twice: as the 'ryr' boundary
of a 12 nm geometry (createdgeometries/G*/G*_nonP_0.h5) and as 2-D masks
(createdgeometries/G*/G*_{nonP_0,inner,outer,uniform}_{10,25}.npy, 0 no
channel, 1 RyR, 2 phosphorylated RyR).  The converter that turns a mask
into the HDF5 geometry its species files need (the 'p_ryr' boundary of
boundaryfluxes.c:476-486, parameters/phosphorylation/*.h5) is not part of
the reference (README.md:28 names it); this module is it.

The masks are 168 x 168 pixels of the geometry's 12 nm voxels, centred on
the 84 x 84 cross-section of the 84^3 geometry: pixel (i, j) is voxel
(i - c0, j - c1) with c = (mask - grid) / 2 per axis, as the shipped G1
pair list and mask show.  A channel goes from the jSR voxel into the cleft
voxel below it, at the planes the base geometry's own 'ryr' boundary uses.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

import csparks as cs


def mask_from_pixels(shape: Sequence[int], ryr: Sequence[Sequence[int]], p_ryr: Sequence[Sequence[int]]) -> np.ndarray:
    m = np.zeros(tuple(shape), dtype=np.int64)
    for v, pts in ((1, ryr), (2, p_ryr)):
        for i, j in pts:
            m[i, j] = v
    return m


def channel_planes(base: cs.Geometry, name: str = "ryr") -> Tuple[int, int]:
    """(from z, to z) of the base geometry's discrete channel boundary."""
    pairs = np.asarray(base.boundary(name).pairs)
    zf, zt = np.unique(pairs[:, 2]), np.unique(pairs[:, 5])
    if len(zf) != 1 or len(zt) != 1:
        raise ValueError(f"the '{name}' boundary of the base geometry spans several planes")
    return int(zf[0]), int(zt[0])


def pattern_geometry(base: cs.Geometry, mask: np.ndarray, channel_boundary: str = "ryr") -> cs.Geometry:
    """The base geometry with its channel boundary replaced by the mask's
    RyRs ('ryr', value 1) and phosphorylated RyRs ('p_ryr', value 2), in
    the mask's row-major pixel order; every other boundary is kept."""
    mask = np.asarray(mask)
    n = base.voxels.shape
    if mask.ndim != 2:
        raise ValueError("the mask must be 2-D")
    off = [(mask.shape[a] - n[a]) // 2 for a in range(2)]
    if any(o < 0 or (mask.shape[a] - n[a]) % 2 for a, o in enumerate(off)):
        raise ValueError(f"mask {mask.shape} is not centred on the {n[:2]} cross-section")
    zf, zt = channel_planes(base, channel_boundary)
    out = {}
    for name, value in (("ryr", 1), ("p_ryr", 2)):
        pix = np.argwhere(mask == value)
        xy = pix - np.asarray(off)
        if len(xy) and (xy.min() < 0 or np.any(xy.max(axis=0) >= np.asarray(n[:2]))):
            raise ValueError(f"a '{name}' pixel of the mask lies outside the geometry")
        pairs = np.zeros((len(xy), 6), dtype=np.uint16)
        pairs[:, 0], pairs[:, 1], pairs[:, 2] = xy[:, 0], xy[:, 1], zf
        pairs[:, 3], pairs[:, 4], pairs[:, 5] = xy[:, 0], xy[:, 1], zt
        for p in pairs:
            if base.voxels[p[0], p[1], p[2]] == base.voxels[p[3], p[4], p[5]]:
                raise ValueError(f"channel at {tuple(p[:2])} does not cross a domain boundary")
        out[name] = pairs
    bnds = [b for b in base.boundaries if b.name not in (channel_boundary, "ryr", "p_ryr")]
    bnds += [cs.Boundary("ryr", "discrete", out["ryr"]), cs.Boundary("p_ryr", "discrete", out["p_ryr"])]
    return cs.Geometry(voxels=base.voxels, h=base.h, domain_names=list(base.domain_names),
                       domain_connections=base.domain_connections, boundaries=bnds, global_size=base.global_size)
