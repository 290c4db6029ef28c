"""Host side of the ensemble workload: the seeded dyad generator (edges,
cluster sizes, phosphorylated fractions, layers, trigger channel)."""
import numpy as np

import dyad


def test_ensemble_dyads_follow_their_distributions():
    ds = [dyad.ensemble_dyad(seed) for seed in range(60)]
    edges, counts, fracs = [], [], []
    for d in ds:
        g = d.geometry
        nx, ny, nz = g.voxels.shape
        assert nx == ny and nz == 25 and g.h == 6.0
        edges.append(nx * g.h)
        ryr, pryr = g.boundary("ryr").pairs, g.boundary("p_ryr").pairs
        n = len(ryr) + len(pryr)
        counts.append(n)
        fracs.append(len(pryr) / n)
        # channels go from the first jSR layer into the last cleft layer
        for p in np.vstack([ryr, pryr]):
            assert g.voxels[p[0], p[1], p[2]] == dyad.JSR and g.voxels[p[3], p[4], p[5]] == dyad.CLEFT
            assert p[2] == p[5] + 1
        # no two channels share a voxel
        xy = {(int(p[0]), int(p[1])) for p in np.vstack([ryr, pryr])}
        assert len(xy) == n
        # the trigger is a non-phosphorylated RyR nearest the cluster centre
        c = nx // 2
        # (pairs are unsigned: signed offsets from the centre)
        dist = [(int(p[0]) - c) ** 2 + (int(p[1]) - c) ** 2 for p in np.vstack([ryr, pryr])]
        t = ryr[d.open_index]
        assert (int(t[0]) - c) ** 2 + (int(t[1]) - c) ** 2 == min(dist)
    assert 256 - 6 <= min(edges) and max(edges) <= 512 + 6
    assert 9 <= min(counts) and max(counts) <= 100
    assert 0.0 <= min(fracs) and max(fracs) <= 0.5 + 1.0 / 9
    # spread over the ranges, not clustered at one end
    assert np.ptp(edges) > 150 and np.ptp(counts) > 50


def test_ensemble_dyads_are_reproducible():
    a, b = dyad.ensemble_dyad(17), dyad.ensemble_dyad(17)
    np.testing.assert_array_equal(a.geometry.voxels, b.geometry.voxels)
    for x, y in zip(a.geometry.boundaries, b.geometry.boundaries):
        np.testing.assert_array_equal(x.pairs, y.pairs)
    assert a.open_index == b.open_index
    c = dyad.ensemble_dyad(18)
    assert c.geometry.voxels.shape != a.geometry.voxels.shape or c.description != a.description
