#!/usr/bin/env python3
"""Time the FP64 C oracle on the shipped G1 geometry (84^3 voxels at
12 nm, 5 species slots, dt 1e-4 ms, RyR 24 forced open until 5 ms), the
run whose main loop the reference times at 2.6 ms per step on 8 OpenMP
threads (its -v timing table, fvm_mpi/src/utils.c:47-75).  Prints the
oracle's ms per step and voxel-steps per second with the thread count
used; run it as

    OMP_NUM_THREADS=8 python3 oracle/time_g1.py [steps]

Test infrastructure: it measures the oracle (the CPU restatement), not
the product, and reads the G1 inputs from tests/golden/."""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path[:0] = [HERE, os.path.join(REPO, "hpc-ca-phospattern_amd"), os.path.join(REPO, "tests")]

import csparks as cs  # noqa: E402
from helpers import load_geometry_fixture, load_model_fixture  # noqa: E402
from oracle_py import Oracle  # noqa: E402


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 1000
    g = load_geometry_fixture("g1_nonP_0")
    m = load_model_fixture("params_sr1100_ryr55")
    orc = Oracle(g, m, cs.Options(h=12.0, dt=1e-4, seed=1, open={"ryr": [24]}, t_close={"ryr": 5.0}))
    nv = 1
    for k in orc.n:
        nv *= int(k)
    orc.step(1)     # first touch of the fields
    t0 = time.perf_counter()
    orc.step(steps)
    el = time.perf_counter() - t0
    orc.close()
    threads = os.environ.get("OMP_NUM_THREADS", "default")
    ms = el * 1e3 / steps
    print(f"G1 {tuple(orc.n)} = {nv} voxels, FP64 oracle, {steps} steps on {threads} threads: "
          f"{ms:.3f} ms/step, {nv * steps / el:.3e} voxel-steps/s ({el / (nv * steps) * 1e9:.2f} ns/voxel-step)")


if __name__ == "__main__":
    main()
