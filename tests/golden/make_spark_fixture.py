#!/usr/bin/env python3
"""Spark statistics of CPU oracle realisations of the fast-gating test dyad
(dyad.fast_spark_dyad), gated with the reference's Mersenne Twister draw
order (mt_seed32new(seed), one mt_ldrand per stochastic channel per gating
step), one realisation per seed.  The GPU runs the same dyad with its own
Philox stream and must produce the same distributions (two-sample KS).

Writes tests/golden/spark_stats_fast_dyad.json (1000 seeds by default,
the linescan sampled every 0.002 ms).  Takes several minutes on 8 cores:
python3 tests/golden/make_spark_fixture.py [num_seeds]
"""
import json
import os
import sys
from multiprocessing import Pool

# one OpenMP thread per worker process: the pool already uses every core
os.environ["OMP_NUM_THREADS"] = "1"

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "hpc-ca-phospattern_amd"), os.path.join(REPO, "oracle")]

import csparks as cs  # noqa: E402
import dyad  # noqa: E402
import sparks  # noqa: E402

# the run every realisation makes (shared with tests/test_gpu_sparks.py)
RUN = dict(tstop_ms=0.3, save_ms=0.002, species=1, domain="cleft")


def plan(d):
    """steps, save interval, linescan offsets of the fixed run."""
    dt = 0.95 * d.geometry.h ** 2 / (6 * max(max(s) for s in d.model.sigma if s))
    nsteps = cs.host_lib().cs_count_steps(RUN["tstop_ms"], dt)
    save = cs.host_lib().cs_save_interval(RUN["save_ms"], dt)
    dom = d.model.domain_names.index(RUN["domain"])
    return nsteps, save, sparks.line_through(d.geometry.voxels, dom)


def one(seed):
    import oracle_py
    d = dyad.fast_spark_dyad()
    nsteps, save, offsets = plan(d)
    orc = oracle_py.Oracle(d.geometry, d.model, cs.Options(seed=seed, rng="mt", open={"ryr": [d.open_index]}))
    assert abs(orc.dt - 0.95 * d.geometry.h ** 2 / (6 * 220000.0)) < 1e-15
    st = sparks.realise(orc, nsteps, save, RUN["species"], offsets, d.model.conv)
    orc.close()
    return {"seed": seed, "peak": st.peak, "duration": st.duration, "recruited": st.recruited}


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 1000
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        rows = pool.map(one, range(1, n + 1))
    d = dyad.fast_spark_dyad()
    nsteps, save, offsets = plan(d)
    out = {"model": "dyad.fast_spark_dyad()", "rng": "mt", "open": {"ryr": [d.open_index]}, **RUN,
           "nsteps": nsteps, "save_every": save, "offsets": list(offsets), "realisations": rows}
    with open(os.path.join(HERE, "spark_stats_fast_dyad.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"{n} realisations written")


if __name__ == "__main__":
    main()
