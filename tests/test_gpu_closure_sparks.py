"""Spark statistics of the 12 nm closure form against the plain 3 nm method.

256 dyads (edges 192-256 nm, 9-30 RyRs, fast spark kinetics) run at 3 nm,
at 12 nm with both closure networks (the committed parameters), and at
12 nm without them.  The closure's peak dF/F0, duration and RyRs
recruited must match the 3 nm distributions (two-sample KS, each at
0.05 / 3), at more than 20 times the 3 nm method's realisations per
second; the plain 12 nm grid must fail at least one statistic, which is
what the closure is for.  Ensemble runs are deterministic (Philox keys,
fixed per-voxel arithmetic), so the outcome is the same on every run."""
import sys

import pytest

import closure_sparks

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_closure_spark_statistics_match_the_fine_method():
    out = closure_sparks.compare(count=256, edge=(192.0, 256.0), ryr=(9, 30), tstop_ms=0.3, say=_say)
    assert out["flux_closure"], "the committed parameters hold no flux network"
    alpha = 0.05 / 3
    bad = [f"{k}: p {out['closure']['ks_vs_fine'][k]['p']:.4f}" for k in closure_sparks.KEYS
           if not out["closure"]["ks_vs_fine"][k]["p"] > alpha]
    assert not bad, "closure vs fine: " + "; ".join(bad)
    assert min(out["coarse"]["ks_vs_fine"][k]["p"] for k in closure_sparks.KEYS) < alpha
    _say(f"closure {out['closure']['realisations_per_s']:.1f} realisations/s, fine "
         f"{out['fine']['realisations_per_s']:.2f}: {out['speedup_closure_vs_fine']}x")
    assert out["speedup_closure_vs_fine"] > 20
