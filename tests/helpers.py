"""Shared test inputs: small dyads the oracle steps in seconds."""
from __future__ import annotations

import json
import os

import numpy as np

import csparks as cs
import dyad


def with_serca(d: dyad.Dyad, density: float = 75.0) -> dyad.Dyad:
    """Add a network-SR slab under the top face and a SERCA density
    boundary from the cytosol layer below it into the SR (the reference's
    'serca' boundary of the shipped geometries, boundaryfluxes.c:12-85)."""
    g = d.geometry
    vox = g.voxels.copy()
    nx, ny, nz = vox.shape
    vox[2:nx - 2, 2:ny - 2, nz - 2:] = dyad.NSR
    pairs = []
    for x in range(2, nx - 2):
        for y in range(2, ny - 2):
            pairs.append((x, y, nz - 3, x, y, nz - 2))
    bnds = list(g.boundaries) + [cs.Boundary("serca", "density", np.array(pairs, dtype=np.uint16))]
    geom = cs.Geometry(voxels=vox, h=g.h, domain_names=list(g.domain_names),
                       domain_connections=g.domain_connections, boundaries=bnds)
    m = d.model
    model = cs.SpeciesModel(domain_names=m.domain_names, species_names=m.species_names,
                            diffusive=m.diffusive, sigma=m.sigma, init=m.init, buffers=m.buffers,
                            ryr=m.ryr, p_ryr=m.p_ryr,
                            serca=cs.SercaModel(307.0, density, 1.0, "serca", "Ca"))
    return dyad.Dyad(geom, model, d.open_index, d.description + " + serca")


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def species_close(a: np.ndarray, b: np.ndarray, rtol: float):
    """Per-voxel relative error, with an absolute floor of rtol times the
    field's magnitude for voxels whose value is (near) zero."""
    scale = float(np.max(np.abs(b)))
    tol = rtol * np.maximum(np.abs(b), 1e-6 * scale)
    bad = np.abs(a - b) > tol
    return not bad.any(), float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6 * max(scale, 1e-300))))


DEV_F32 = 1e-4   # FP32 fields against FP64: error per species over its largest deviation from rest


def rest_field(d: dyad.Dyad, s: int) -> np.ndarray:
    """Resting value of species s per voxel of the dyad's own grid (the
    species file's init value of the voxel's domain)."""
    return np.asarray(d.model.init)[d.geometry.voxels, s]


def deviation_error(got: np.ndarray, ref: np.ndarray, rest: np.ndarray):
    """(max |got - ref|, max |ref - rest|), rest per voxel."""
    dev = float(np.max(np.abs(ref - rest)))
    err = float(np.max(np.abs(got - ref)))
    return err, dev


def deviation_close(got: np.ndarray, ref: np.ndarray, rest: np.ndarray, rtol: float):
    """FP32 fields hold deviations from rest, so their error scales with the
    deviation, not with the value: |got - ref| <= rtol x max |ref - rest|,
    plus 1e-12 of the resting value for a species that has not moved."""
    err, dev = deviation_error(got, ref, rest)
    floor = 1e-12 * float(np.max(np.abs(rest)))
    return err <= rtol * dev + floor, err / max(dev, 1e-300), err, dev


def snap_to_atoms(got: np.ndarray, ref: np.ndarray, rtol: float = 1e-4):
    """A statistic that the reference sample repeats exactly comes from one
    deterministic trajectory (a forced-open channel that stays open for the
    whole window and recruits no other), an atom of its distribution.  FP32
    fields put that trajectory a few ulps away from FP64, and a two-sample
    Kolmogorov-Smirnov test between samples whose atoms sit at slightly
    different points reports the atom's whole mass as D.  Values of `got`
    within rtol of an atom are moved onto it; rtol 1e-4 lies well above the
    FP32 rounding of a trajectory (~1e-6) and below the gap between an atom
    and the nearest run that ends differently (a channel closing just
    before the window ends, ~3e-4 in the lattice fixture).  Returns
    (snapped copy, largest relative move, number moved, atoms)."""
    vals, counts = np.unique(ref, return_counts=True)
    atoms = vals[counts > 1]
    out = np.array(got, dtype=np.float64)
    moved, worst = 0, 0.0
    for v in atoms:
        near = (np.abs(out - v) <= rtol * abs(v)) & (out != v)
        if near.any():
            moved += int(near.sum())
            worst = max(worst, float(np.max(np.abs(out[near] - v))) / abs(v))
            out[near] = v
    return out, worst, moved, atoms


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_geometry_fixture(name: str) -> cs.Geometry:
    """A reference geometry extracted by tests/golden/make_fixtures.py."""
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    bnds = [cs.Boundary(b["name"], b["type"], z["boundary_" + b["name"]]) for b in meta["boundaries"]]
    return cs.Geometry(voxels=z["voxels"], h=meta["h"], domain_names=meta["domain_names"],
                       domain_connections=z["domain_connections"], boundaries=bnds,
                       global_size=tuple(meta["global_size"]))


def load_model_fixture(name: str) -> cs.SpeciesModel:
    d = json.load(open(os.path.join(GOLDEN, name + ".json")))
    ch = lambda c: cs.ChannelModel(**c) if c else None
    return cs.SpeciesModel(domain_names=d["domain_names"], species_names=d["species_names"],
                           diffusive=d["diffusive"], sigma=d["sigma"], init=np.array(d["init"]),
                           buffers=[[cs.Buffer(**b) for b in dom] for dom in d["buffers"]],
                           ryr=ch(d["ryr"]), p_ryr=ch(d["p_ryr"]),
                           serca=cs.SercaModel(**d["serca"]) if d["serca"] else None,
                           conv=tuple(d["conv"]), fixed=[tuple(x) for x in d["fixed"]])
