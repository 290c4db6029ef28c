"""Spark statistics of GPU realisations against the oracle's.

The oracle realisations (tests/golden/spark_stats_fast_dyad.json: a
fast-gating dyad, 1000 seeds; tests/golden/spark_stats_lattice_dyad.json:
the lattice dyad with the shipped kinetics, 200 seeds) gate with the
reference's Mersenne Twister order in FP64; the GPU ones use the Philox
stream with FP32 fields.  The random streams differ, so the
realisations differ one by one; the distributions of peak dF/F0, duration
and RyRs recruited must agree (two-sample Kolmogorov-Smirnov, p > 0.05).
Values the oracle repeats exactly (one deterministic trajectory, e.g. the
forced-open channel open for the whole window) are atoms; GPU values
within 1e-4 of one are the same outcome in FP32 and are put on it before
the test (helpers.snap_to_atoms)."""
import json
import os
import sys

import numpy as np
import pytest
from scipy.stats import ks_2samp

import csparks as cs
import dyad
import sparks
from conftest import REPO
from helpers import snap_to_atoms

pytestmark = pytest.mark.gpu
FIXTURE = os.path.join(REPO, "tests", "golden", "spark_stats_fast_dyad.json")
LATTICE = os.path.join(REPO, "tests", "golden", "spark_stats_lattice_dyad.json")


def _compare(fx, d, label):
    ref = fx["realisations"]
    n = len(ref)
    print(f"{label}: {n} GPU members (Philox, FP32) against {n} oracle seeds (MT, FP64)", flush=True)
    ens = cs.Ensemble([(d.geometry, d.model, cs.Options(seed=s, precision="f32", open={"ryr": fx["open"]["ryr"]}))
                       for s in range(1, n + 1)], instances=list(range(n)))
    got = sparks.realise_ensemble(ens, fx["nsteps"], fx["save_every"], fx["species"], fx["offsets"], d.model.conv)
    ens.close()
    alpha = 0.05 / 3
    failed = []
    for key in ("peak", "duration", "recruited"):
        b = np.array([r[key] for r in ref])
        a, worst, moved, atoms = snap_to_atoms(np.array([getattr(g, key) for g in got]), b)
        res = ks_2samp(a, b)
        print(f"{key}: GPU mean {a.mean():.4g}, oracle mean {b.mean():.4g}, KS D {res.statistic:.4f} "
              f"p {res.pvalue:.4f}; {moved} GPU values snapped onto {len(atoms)} oracle atoms "
              f"(largest move {worst:.1e})", flush=True)
        if not res.pvalue > alpha:
            failed.append(f"{key}: KS p = {res.pvalue:.4f}")
    assert not failed, "; ".join(failed)


def test_spark_statistics_match_the_oracle_distributions():
    """All realisations run as one ensemble (one member per seed).  Three
    statistics are tested at once, so each must pass at 0.05 / 3
    (Bonferroni), which keeps the family's false-alarm rate under 5%."""
    with open(FIXTURE) as f:
        fx = json.load(f)
    assert len(fx["realisations"]) >= 1000
    _compare(fx, dyad.fast_spark_dyad(), "fast-gating dyad")


@pytest.mark.skipif(not os.path.exists(LATTICE), reason="spark_stats_lattice_dyad.json not generated "
                    "(tests/golden/make_spark_fixture.py lattice)")
def test_lattice_spark_statistics_with_the_shipped_kinetics():
    """The 64x64x25 lattice dyad with the shipped RyR kinetics (Hill
    opening at K 70 / 35 uM up to 0.3/ms, closing 0.5/ms), the first 2 ms
    of 200 sparks (tests/golden/spark_stats_lattice_dyad.json)."""
    with open(LATTICE) as f:
        fx = json.load(f)
    assert len(fx["realisations"]) >= 200
    _compare(fx, dyad.lattice_dyad(), "lattice dyad, shipped kinetics")
