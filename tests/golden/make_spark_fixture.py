#!/usr/bin/env python3
"""Spark statistics of CPU oracle realisations, gated with the reference's
Mersenne Twister draw order (mt_seed32new(seed), one mt_ldrand per
stochastic channel per gating step), one realisation per seed.  The GPU
runs the same dyad with its own Philox stream and must produce the same
distributions (two-sample KS).

Workloads:
  fast     dyad.fast_spark_dyad(): RyR kinetics several hundred times the
           shipped ones, 0.3 ms, linescan every 0.002 ms; 1000 seeds by
           default -> spark_stats_fast_dyad.json (several minutes on 8 cores);
  lattice  dyad.lattice_dyad(): the 64x64x25 lattice dyad with the shipped
           kinetics (Hill opening, K 70 / 35 uM, 0.3/ms, closing 0.5/ms),
           the first 2 ms of each spark, linescan every 0.02 ms; 200 seeds
           by default -> spark_stats_lattice_dyad.json (about 1.5 hours on
           8 cores: 77,000 steps of 102,400 voxels per seed).

python3 tests/golden/make_spark_fixture.py [fast|lattice] [num_seeds] [workers]
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
WORKLOADS = {
    "fast": dict(make=lambda: dyad.fast_spark_dyad(), model="dyad.fast_spark_dyad()", seeds=1000,
                 out="spark_stats_fast_dyad.json",
                 run=dict(tstop_ms=0.3, save_ms=0.002, species=1, domain="cleft")),
    "lattice": dict(make=lambda: dyad.lattice_dyad(), model="dyad.lattice_dyad()", seeds=200,
                    out="spark_stats_lattice_dyad.json",
                    run=dict(tstop_ms=2.0, save_ms=0.02, species=1, domain="cleft")),
}
WORKLOAD = "fast"
RUN = WORKLOADS[WORKLOAD]["run"]


def plan(d):
    """steps, save interval, linescan offsets of the fixed run."""
    dt = 0.95 * d.geometry.h ** 2 / (6 * max(max(s) for s in d.model.sigma if s))
    nsteps = cs.host_lib().cs_count_steps(RUN["tstop_ms"], dt)
    save = cs.host_lib().cs_save_interval(RUN["save_ms"], dt)
    dom = d.model.domain_names.index(RUN["domain"])
    return nsteps, save, sparks.line_through(d.geometry.voxels, dom)


def one(job):
    import oracle_py
    global RUN
    workload, seed = job
    RUN = WORKLOADS[workload]["run"]
    d = WORKLOADS[workload]["make"]()
    nsteps, save, offsets = plan(d)
    orc = oracle_py.Oracle(d.geometry, d.model, cs.Options(seed=seed, rng="mt", open={"ryr": [d.open_index]}))
    assert abs(orc.dt - 0.95 * d.geometry.h ** 2 / (6 * 220000.0)) < 1e-15
    st = sparks.realise(orc, nsteps, save, RUN["species"], offsets, d.model.conv)
    orc.close()
    return {"seed": seed, "peak": st.peak, "duration": st.duration, "recruited": st.recruited}


def main():
    global RUN
    workload = sys.argv[1] if len(sys.argv) > 1 else "fast"
    w = WORKLOADS[workload]
    RUN = w["run"]
    n = int(sys.argv[2]) if len(sys.argv) > 2 else w["seeds"]
    workers = int(sys.argv[3]) if len(sys.argv) > 3 else min(8, os.cpu_count() or 1)
    with Pool(workers) as pool:
        rows = pool.map(one, [(workload, s) for s in range(1, n + 1)], chunksize=1)
    d = w["make"]()
    nsteps, save, offsets = plan(d)
    out = {"model": w["model"], "rng": "mt", "open": {"ryr": [d.open_index]}, **RUN,
           "nsteps": nsteps, "save_every": save, "offsets": list(offsets), "realisations": rows}
    with open(os.path.join(HERE, w["out"]), "w") as f:
        json.dump(out, f, indent=0)
    print(f"{n} realisations written to {w['out']}")


if __name__ == "__main__":
    main()
