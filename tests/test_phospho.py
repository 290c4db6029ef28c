"""The phosphorylation-pattern converter (phospho.py) against the files
the reference ships for G1: its 'nonP_0' mask reproduces the 50 'ryr'
pairs of G1_nonP_0.h5 exactly, every pattern phosphorylates a subset of
those same 50 channels with the advertised counts, the inner and outer
patterns put the pRyRs at the centre and the rim of the cluster, and the
converted geometry round-trips through the HDF5 writer and builds with
the phosphorylation species file."""
import json
import os

import numpy as np
import pytest

import csparks as cs
import phospho
from helpers import GOLDEN, load_geometry_fixture, load_model_fixture

MASKS = json.load(open(os.path.join(GOLDEN, "g1_phospho_masks.json")))


def _mask(name):
    m = MASKS[name]
    return phospho.mask_from_pixels(m["shape"], m["ryr"], m["p_ryr"])


def test_nonP_mask_reproduces_the_shipped_ryr_pairs():
    base = load_geometry_fixture("g1_nonP_0")
    g = phospho.pattern_geometry(base, _mask("G1_nonP_0"))
    np.testing.assert_array_equal(g.boundary("ryr").pairs, base.boundary("ryr").pairs)
    assert len(g.boundary("p_ryr").pairs) == 0
    assert [b.name for b in g.boundaries if b.name not in ("ryr", "p_ryr")] == ["serca", "sr_cleft"]


@pytest.mark.parametrize("name,npryr", [("G1_inner_10", 10), ("G1_inner_25", 25), ("G1_outer_10", 10),
                                        ("G1_outer_25", 25), ("G1_uniform_10", 10), ("G1_uniform_25", 25)])
def test_patterns_phosphorylate_a_subset_of_the_same_channels(name, npryr):
    base = load_geometry_fixture("g1_nonP_0")
    g = phospho.pattern_geometry(base, _mask(name))
    r, p = g.boundary("ryr").pairs, g.boundary("p_ryr").pairs
    assert len(p) == npryr and len(r) == 50 - npryr
    together = sorted(map(tuple, np.vstack([r, p]).tolist()))
    assert together == sorted(map(tuple, base.boundary("ryr").pairs.tolist()))


def test_inner_and_outer_patterns_sit_at_the_centre_and_the_rim():
    base = load_geometry_fixture("g1_nonP_0")
    centre = base.boundary("ryr").pairs[:, :2].astype(float).mean(axis=0)
    def mean_dist(name, which):
        pts = phospho.pattern_geometry(base, _mask(name)).boundary(which).pairs[:, :2].astype(float)
        return float(np.mean(np.linalg.norm(pts - centre, axis=1)))
    for k in ("10", "25"):
        assert mean_dist(f"G1_inner_{k}", "p_ryr") < mean_dist(f"G1_inner_{k}", "ryr")
        assert mean_dist(f"G1_outer_{k}", "p_ryr") > mean_dist(f"G1_outer_{k}", "ryr")


def test_converted_geometry_round_trips_and_builds(tmp_path):
    base = load_geometry_fixture("g1_nonP_0")
    g = phospho.pattern_geometry(base, _mask("G1_inner_10"))
    path = str(tmp_path / "G1_inner_10.h5")
    cs.write_geometry(path, g)
    back = cs.read_geometry(path)
    for name in ("ryr", "p_ryr", "serca", "sr_cleft"):
        np.testing.assert_array_equal(back.boundary(name).pairs, g.boundary(name).pairs)
    np.testing.assert_array_equal(back.voxels, g.voxels)
    model = load_model_fixture("params_sr900_ryr55_pryr25")
    s = cs.Setup(back, model, cs.Options(h=12.0, seed=1, open={"ryr": [0]}, t_close={"ryr": 1.0}))
    kinds = [(int(s.problem.fluxes[f].num_pairs), int(s.problem.fluxes[f].discrete))
             for f in range(s.problem.num_fluxes)]
    assert (40, 1) in kinds and (10, 1) in kinds
    s.close()


def test_converter_rejects_masks_that_do_not_fit():
    base = load_geometry_fixture("g1_nonP_0")
    with pytest.raises(ValueError):
        phospho.pattern_geometry(base, np.zeros((167, 168), dtype=int))
    bad = np.zeros((168, 168), dtype=int)
    bad[0, 0] = 1            # outside the central 84 x 84 crop
    with pytest.raises(ValueError):
        phospho.pattern_geometry(base, bad)
