#!/usr/bin/env python3
"""Benchmark: calcium_sparks time stepping on one MI355X.

Headline workload (the single-instance configuration): a 1200x1200x600 nm
dyad at 3 nm (400x400x200 = 32M voxels), 5 species slots (7 species),
cleft 12 nm, jSR 30 nm over 600x600 nm, 100 RyRs by seeded Poisson-disk
sampling in a 400 nm disk (30% pRyR), stochastic Philox gating, FP32
fields.  A step is one full time step of the scheme (gating, stencil +
reaction, boundary fluxes).  Inputs are synthetic and resident in HBM
before timing starts.

Prints one JSON line.  `value` is the whole-step algorithmic HBM
bandwidth: 41 B per voxel-step (1 B domain id + 5 species x (4 B read +
4 B write)) x voxels / step time.  `roofline` prices the stencil pass
alone (HIP events on its stream): a one-step pass (k_stencil) or a fused
two-step pass (k_step2, the default where the problem admits it), priced
at 41 B per voxel-step times the voxel-steps one launch processes (one
or two per voxel).  `traffic` is the HBM bytes per launch that committed
rocprofv3 counters measured for the same launch shape and kernel source
(null when the kernel source changed since).  `cpu_baseline` runs the
FP64 C oracle on the host cores for a bounded sample of the same workload.

Sub-records of the same line:
  fp64      the same dyad with FP64 fields (one-step passes), the
            reference's own precision, and the per-species error of the
            FP32 fields against the FP64 ones after 200 steps;
  g1        the run the reference publishes its own timing for (G1 at 12
            nm, 1000 steps, 2.6 ms per step on 8 CPU threads), per step on
            the GPU in FP64 and FP32;
  ensemble  512 realisations of 256-512 nm dyads at 6 nm (9-100 RyRs,
            0-50% pRyR) stepped as one ensemble: spark realisations/s over
            their 20 ms runs;
  closure   realisations/s of 4096 dyads of 512-1024 nm (25-150 RyRs) at 12
            nm with the learned closure (gating and flux networks), against
            the plain ensemble of 16 of the same dyads at 3 nm, over their
            50 ms runs; then the closure's spark statistics against the 3 nm
            method's on 1024 small dyads (--ks-count) by two-sample KS.
Realisations/s = realisations / (steps of one run x measured step time).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "hpc-ca-phospattern_amd"))

import numpy as np  # noqa: E402

import csparks as cs  # noqa: E402
import dyad  # noqa: E402
import hashlib  # noqa: E402

PEAK_HBM_GBS = 8000.0
BYTES_F32 = 1 + 5 * (4 + 4)
BYTES_F64 = 1 + 5 * (8 + 8)
METRIC = "spark realisations/sec (closure vs plain) at KS<0.05; stencil HBM GB/s %peak"
TSTOP_MS = 30.0
PROFILE_TRAFFIC = os.path.join(REPO, "profiles", "traffic.json")
WORKLOAD_KEY = "cluster_dyad"
CLOSURE_PARAMS = os.path.join(REPO, "hpc-ca-phospattern_amd", "closure_params.npz")
LAUNCH_KEYS = ["rows", "zchunk", "xcd_remap", "prefetch", "nontemporal", "lds_tile", "linear", "yblock",
               "fuse", "fuse_zchunk", "fuse_xcd", "fuse_fixed"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(d, budget_s=15.0):
    """FP64 oracle on the host cores, a bounded number of steps of the same
    grid; returns voxel-steps/s."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py as op
    opts = cs.Options(h=3.0, seed=1, open={"ryr": [0]})
    orc = op.Oracle(d.geometry, d.model, opts)
    nv = int(np.prod(orc.n))
    orc.step(1)  # first touch of the fields
    t0 = time.perf_counter()
    steps = 0
    while steps < 3 or (time.perf_counter() - t0 < budget_s and steps < 200):
        orc.step(1)
        steps += 1
    el = time.perf_counter() - t0
    orc.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return nv * steps / el, cores, steps, el


TRAFFIC_SOURCES = ("csparks.hip", "cs_device.h", "cs_physics.h")


def kernel_source_hash() -> str:
    """SHA-256 of the HIP sources of the single-domain kernels (k_stencil,
    k_step2 and what they include), which the committed traffic profile
    records: a profile of other sources is stale."""
    h = hashlib.sha256()
    csrc = os.path.join(REPO, "hpc-ca-phospattern_amd", "csrc")
    for name in TRAFFIC_SOURCES:
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()


def fp64_record(size, steps):
    """The headline dyad with FP64 fields (one-step passes)."""
    d = dyad.cluster_dyad(seed=1, n=tuple(size))
    sim = cs.Simulation(d.geometry, d.model, cs.Options(h=3.0, seed=1, open={"ryr": [0]}, precision="f64"))
    nv = int(np.prod(sim.n))
    sim.step(5)
    sim.synchronize()
    t1 = time.perf_counter()
    total_ms, pass_ms, passes, fused = sim.time_passes(steps)
    sim.synchronize()
    wall = time.perf_counter() - t1
    sim.close()
    ms = wall * 1e3 / steps
    kern = (steps / passes) * nv * BYTES_F64 / (pass_ms * 1e-3) / 1e9
    log(f"fp64: {ms:.4f} ms/step, pass {pass_ms * 1e3:.1f} us ({kern:.0f} GB/s)")
    return {"ms_per_step": round(ms, 4), "value": round(nv * BYTES_F64 / (ms * 1e-3) / 1e9, 1), "unit": "GB/s",
            "bytes_per_voxel_step": BYTES_F64, "kernel": "k_stencil" if fused == 0 else "k_step2",
            "kernel_ms": round(pass_ms, 5),
            "roofline": {"bound": "hbm", "achieved": round(kern, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(kern / PEAK_HBM_GBS, 4)}}


def fp32_error_record(size, steps=200):
    """Per-species error of the FP32 fields against FP64 ones on the same
    dyad after `steps` steps from rest (same seed and channel trajectories;
    FP32 in its default fused passes, FP64 in one-step passes): the largest
    |f32 - f64| over the grid divided by the largest deviation of the FP64
    field from rest.  The FP64 device path equals the FP64 oracle to 1e-10
    in the parity tests."""
    d = dyad.cluster_dyad(seed=1, n=tuple(size))
    sims = {p: cs.Simulation(d.geometry, d.model, cs.Options(h=3.0, seed=1, open={"ryr": [0]}, precision=p))
            for p in ("f64", "f32")}
    s64, s32 = sims["f64"], sims["f32"]
    ns = s64.problem.num_species
    rest = [s64.species(s) for s in range(ns)]
    for sim in (s64, s32):
        sim.step(steps)
        sim.synchronize()
    same = all(np.array_equal(s64.channel_states(f), s32.channel_states(f)) for f in range(s64.problem.num_fluxes)
               if s64.problem.fluxes[f].discrete)
    errs, devs = [], []
    for s in range(ns):
        a, b = s64.species(s), s32.species(s)
        dev = float(np.max(np.abs(a - rest[s])))
        err = float(np.max(np.abs(b - a)))
        devs.append(dev)
        errs.append(err / dev if dev > 0 else err)
        del a, b
    for sim in (s64, s32):
        sim.close()
    log("fp32 vs fp64 after %d steps, error / max deviation from rest: %s" % (steps, ", ".join(f"{e:.2e}" for e in errs)))
    return {"steps": steps, "error_over_max_deviation": [float(f"{e:.3e}") for e in errs],
            "max_deviation_uM": [float(f"{v:.4g}") for v in devs], "same_channel_trajectories": bool(same)}


G1_REFERENCE_MS = 2.6   # the reference's main-loop time per step on this run, 8 OpenMP threads


def g1_inputs():
    """The reference's G1 geometry (G1_nonP_0.h5, 84^3 at 12 nm) and its
    sr1100/ryr55 species file, as the committed fixtures hold them (read by
    the tests' fixture loaders, which use only the package)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import load_geometry_fixture, load_model_fixture
    return load_geometry_fixture("g1_nonP_0"), load_model_fixture("params_sr1100_ryr55")


def g1_record(steps=1000):
    """The run the reference publishes its own timing for: G1 at 12 nm, 5
    species slots, dt 1e-4 ms, RyR 24 forced open until 5 ms, 1000 steps;
    wall time per step on one GPU (gating, stencil, fluxes, SERCA) in FP64
    and FP32, against the reference's 2.6 ms per step on 8 CPU threads."""
    g, m = g1_inputs()
    out = {"workload": "G1 84^3 at 12 nm (592,704 voxels, 31,416 SERCA pairs, 50 RyRs), sr1100/ryr55 species, "
                       "dt 1e-4 ms, -O ryr 24 -C ryr 5, 1000 steps",
           "reference_cpu_ms_per_step": G1_REFERENCE_MS, "reference_cpu": "1 MPI rank x 8 OpenMP threads"}
    for p in ("f64", "f32"):
        sim = cs.Simulation(g, m, cs.Options(h=12.0, dt=1e-4, seed=1, open={"ryr": [24]}, t_close={"ryr": 5.0},
                                             precision=p))
        sim.step(20)
        sim.synchronize()
        t1 = time.perf_counter()
        _, pass_ms, passes, fused = sim.time_passes(steps)
        sim.synchronize()
        ms = (time.perf_counter() - t1) * 1e3 / steps
        sim.close()
        out[p] = {"ms_per_step": round(ms, 5), "pass_us": round(pass_ms * 1e3, 2), "fused_passes": fused,
                  "passes": passes, "vs_reference_cpu": round(G1_REFERENCE_MS / ms, 1)}
        log(f"G1 {p}: {ms * 1e3:.1f} us/step ({fused}/{passes} fused passes, pass {pass_ms * 1e3:.1f} us), "
            f"{G1_REFERENCE_MS / ms:.0f}x the reference's 2.6 ms/step")
    return out


def ensemble_members(count, h, edge, ryr, seed0, precision, kinetics=None):
    ds = [dyad.ensemble_dyad(seed0 + k, h=h, edge_nm=edge, ryr=ryr) for k in range(count)]
    return ds, [(d.geometry, d.model if kinetics is None else kinetics(d.model),
                 cs.Options(h=h, seed=seed0 + k, open={"ryr": [d.open_index]}, precision=precision))
                for k, d in enumerate(ds)]


def tuned(ens):
    """The ensemble's autotuned launch shape, with its step periods replayed
    as captured graphs; returns the shape."""
    best, _ = ens.autotune(2)
    ens.set_launch(**best)
    ens.set_graph(True)
    return best


def ensemble_rate(ens, steps, tstop_ms, warmup=10):
    """Measured step time of an ensemble and its realisations/s over runs
    of tstop_ms.  With graphs, warmup and steps are whole periods, so the
    warmup captures the graphs and the timed steps only replay them."""
    ens.step(warmup)
    ens.synchronize()
    t1 = time.perf_counter()
    ens.step(steps)
    ens.synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / steps
    run_steps = cs.host_lib().cs_count_steps(tstop_ms, ens.dt)
    return ms, run_steps, ens.count / (run_steps * ms * 1e-3)


def ensemble_record(count=512, steps=100):
    """Config-2 ensemble: realisations of 256-512 nm dyads at 6 nm."""
    t0 = time.perf_counter()
    ds, members = ensemble_members(count, 6.0, (256.0, 512.0), (9, 100), 1, "f32")
    ens = cs.Ensemble(members, instances=list(range(count)))
    _, nv, nbytes = ens.info()
    launch = tuned(ens)
    P = ens.period
    log(f"ensemble: {count} dyads, {nv} voxels, set up and tuned in {time.perf_counter() - t0:.1f}s: {launch}, "
        f"graph period {P}")
    ms, run_steps, rate = ensemble_rate(ens, P * max(1, steps // P), 20.0, warmup=P)
    ens.set_graph(False)
    tot_ms, stencil_ms = ens.time_steps(steps)
    ens.close()
    gbs = nv * BYTES_F32 / (stencil_ms * 1e-3) / 1e9
    log(f"ensemble: {ms:.4f} ms/step (stencil {stencil_ms * 1e3:.1f} us, {gbs:.0f} GB/s) -> {rate:.3f} realisations/s")
    return {"workload": f"{count} dyads, edges 256-512 nm at 6 nm, 9-100 RyRs, 0-50% pRyR, Philox, FP32, 20 ms runs",
            "realisations": count, "voxels": nv, "ms_per_step": round(ms, 4), "steps_per_realisation": run_steps,
            "realisations_per_s": round(rate, 4), "launch": launch, "graphs": True,
            "stencil": {"kernel": "k_batch_stencil" if launch.get("persistent") else "k_batch_tiles",
                        "kernel_ms": round(stencil_ms, 5),
                        "achieved": round(gbs, 1), "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4)}}


def closure_record(count=4096, fine_count=16, steps=120, fine_steps=40, ks=True, ks_count=1024):
    """Config-4 comparison: the 12 nm closure ensemble (both networks)
    against the plain 3 nm ensemble of the same dyads (edges 512-1024 nm,
    25-150 RyRs), step rates at that size; then, with ks, the spark
    statistics of the closure against the 3 nm method on ks_count small
    dyads (closure_sparks.compare), the KS test the metric is quoted at."""
    import closure
    t0 = time.perf_counter()
    net, norm, trained = closure.load(CLOSURE_PARAMS)
    flux = closure.load_flux(CLOSURE_PARAMS)
    ds, members = ensemble_members(count, 12.0, (512.0, 1024.0), (25, 150), 5000, "f32")
    ens = cs.Ensemble(members, instances=list(range(count)))
    ens.set_gating(cs.Ensemble.GATE_CLOSURE, net)
    if flux is not None:
        ens.set_flux_closure(flux[0])
    _, nv12, _ = ens.info()
    launch12 = tuned(ens)
    P = ens.period
    log(f"closure 12 nm: set up and tuned in {time.perf_counter() - t0:.1f}s: {launch12}, graph period {P}")
    ms12, steps12, rate12 = ensemble_rate(ens, P * max(1, steps // P), 50.0, warmup=P)
    flagged = int(np.count_nonzero(ens.monitor()))
    ens.close()
    log(f"closure 12 nm: {count} dyads, {nv12} voxels, {ms12:.4f} ms/step -> {rate12:.3f} realisations/s "
        f"({flagged} flagged by the range monitor)")
    fine = [(g, m, cs.Options(h=3.0, seed=o.seed, open=o.open, precision="f32")) for g, m, o in members[:fine_count]]
    ens3 = cs.Ensemble(fine, instances=list(range(fine_count)))
    _, nv3, _ = ens3.info()
    launch3 = tuned(ens3)
    P3 = ens3.period
    ms3, steps3, rate3 = ensemble_rate(ens3, P3 * max(1, fine_steps // P3), 50.0, warmup=P3)
    ens3.close()
    net.close()
    if flux is not None:
        flux[0].close()
    log(f"plain 3 nm: {fine_count} dyads, {nv3} voxels, {ms3:.4f} ms/step -> {rate3:.5f} realisations/s "
        f"(set up in {time.perf_counter() - t0:.1f}s)")
    rec = {"workload": "dyads of 512-1024 nm edges, 25-150 RyRs, 0-50% pRyR, Philox, FP32, 50 ms runs",
            "closure": {"h_nm": 12.0, "realisations": count, "voxels": nv12, "ms_per_step": round(ms12, 4),
                        "steps_per_realisation": steps12, "realisations_per_s": round(rate12, 4),
                        "network": "trained" if trained else "untrained (same cost)",
                        "flux_closure": flux is not None,
                        "range_monitor_flagged": flagged, "launch": launch12, "graphs": True},
            "plain": {"h_nm": 3.0, "realisations": fine_count, "voxels": nv3, "ms_per_step": round(ms3, 4),
                      "steps_per_realisation": steps3, "realisations_per_s": round(rate3, 6), "launch": launch3,
                      "graphs": True},
            "speedup": round(rate12 / rate3, 1)}
    if ks:
        import closure_sparks
        t1 = time.perf_counter()
        out = closure_sparks.compare(count=ks_count, edge=(192.0, 256.0), ryr=(9, 30), tstop_ms=0.3, say=log)
        rec["ks"] = {"workload": f"{ks_count} dyads of 192-256 nm edges, 9-30 RyRs, fast spark kinetics, 0.3 ms: 12 nm "
                                 "closure vs the plain 3 nm method on the same dyads (two-sample KS)",
                     "closure_vs_fine": {k: {"D": round(v["D"], 4), "p": round(v["p"], 4)}
                                         for k, v in out["closure"]["ks_vs_fine"].items()},
                     "plain_12nm_vs_fine": {k: {"D": round(v["D"], 4), "p": round(v["p"], 6)}
                                            for k, v in out["coarse"]["ks_vs_fine"].items()},
                     "all_p_above_0.05": all(v["p"] > 0.05 for v in out["closure"]["ks_vs_fine"].values()),
                     "flagged_to_fine_fallback": out["flagged"],
                     "speedup_at_this_workload": out["speedup_closure_vs_fine"],
                     "seconds": round(time.perf_counter() - t1, 1)}
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--size", type=int, nargs=3, default=[400, 400, 200])
    ap.add_argument("--autotune", action="store_true",
                    help="time the candidate launch shapes first and keep the fastest (default: the library's "
                         "default shape, which the committed profiles cover)")
    ap.add_argument("--launch", type=str, default="",
                    help="fixed launch shape " + ",".join(LAUNCH_KEYS) + " (a prefix of them; skips autotune)")
    ap.add_argument("--no-extra", action="store_true", help="only the headline dyad (no fp64, ensemble, closure)")
    ap.add_argument("--no-ks", action="store_true", help="skip the closure's spark-statistics KS comparison")
    ap.add_argument("--ks-count", type=int, default=1024,
                    help="realisations of the KS comparison (each runs at 12 nm and at 3 nm)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (for scaling traffic counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        ms = cs.calibration_copy(args.copy_calibration << 20, max(args.steps, 1))
        print(json.dumps({"copy_bytes": args.copy_calibration << 20, "ms_per_copy": ms,
                          "GBps": 2 * (args.copy_calibration << 20) / (ms * 1e-3) / 1e9}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported: the domain stays "
              "resident on one GPU (independent realisations would shard by instance)", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    d = dyad.cluster_dyad(seed=1, n=tuple(args.size))
    opts = cs.Options(h=3.0, seed=1, open={"ryr": [0]}, precision=args.precision, tstop=TSTOP_MS)
    sim = cs.Simulation(d.geometry, d.model, opts)
    log(f"setup {time.perf_counter() - t0:.1f}s; {cs.device_info()}; grid {sim.n}")
    nv = int(np.prod(sim.n))
    bpv = BYTES_F32 if args.precision == "f32" else BYTES_F64

    launch = sim.launch_config()
    if args.launch:
        sim.set_launch_config(**dict(zip(LAUNCH_KEYS, (int(v) for v in args.launch.split(",")))))
        launch = sim.launch_config()
    sim.step(args.warmup)
    sim.synchronize()
    if args.autotune and not args.launch:
        t2 = time.perf_counter()
        launch, tune_ms = sim.autotune(1)
        log(f"autotune {time.perf_counter() - t2:.1f}s -> {launch} ({tune_ms * 1e3:.1f} us)")
    t1 = time.perf_counter()
    total_ms, kernel_ms, passes, fused = sim.time_passes(args.steps)
    sim.synchronize()
    wall = time.perf_counter() - t1
    ms_per_step = wall * 1e3 / args.steps
    gbs = nv * bpv / (ms_per_step * 1e-3) / 1e9
    steps_per_launch = args.steps / passes
    kern_gbs = steps_per_launch * nv * bpv / (kernel_ms * 1e-3) / 1e9
    steps_per_real = cs.host_lib().cs_count_steps(TSTOP_MS, sim.dt)
    kernel = "k_step2" if fused == passes else "k_stencil" if fused == 0 else "k_step2+k_stencil"
    log(f"{args.steps} steps in {passes} passes ({fused} fused): wall {wall * 1e3:.1f} ms, "
        f"events {total_ms:.1f} ms, pass {kernel_ms * 1e3:.1f} us ({kern_gbs:.0f} GB/s)")

    # conservation check of the timed run (total calcium, every slot)
    tot = sim.stats()[2].sum()
    sim.close()

    # HBM bytes per launch of this kernel and launch shape, from the
    # committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by the copy
    # kernel's readings); null when no profile matches
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            t = json.load(f)
        key = ",".join(str(launch[k]) for k in LAUNCH_KEYS)
        if t.get("workload") == WORKLOAD_KEY and t.get("precision") == args.precision and \
                tuple(t.get("grid", ())) == tuple(args.size) and t.get("source_sha256") == kernel_source_hash():
            traffic = t.get("bytes_per_launch", {}).get(key)

    cpu = None
    if not args.no_cpu:
        vps, cores, csteps, cel = cpu_baseline(d)
        cpu = {"value": round(vps * BYTES_F32 / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "port",
               "sample": f"FP64 C oracle, {csteps} steps of the same {args.size[0]}x{args.size[1]}x{args.size[2]} grid "
                         f"in {cel:.1f} s ({vps:.3e} voxel-steps/s, priced at {BYTES_F32} B/voxel-step)"}

    out = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "f32" else "f64",
        "data": "synthetic",
        "config": {"workload": f"single {args.size[0]}x{args.size[1]}x{args.size[2]} dyad @3nm, 7 species in "
                               "5 slots, 100 RyR (30% pRyR), stochastic Philox gating, "
                               f"{'FP32' if args.precision == 'f32' else 'FP64'} fields",
                   "grid": list(args.size), "voxels": nv,
                   "bytes_per_voxel_step": bpv},
        "frac_of_peak": round(gbs / PEAK_HBM_GBS, 4),
        "voxel_steps_per_s": round(nv / (ms_per_step * 1e-3), 1),
        "spark_realisations_per_s": 1.0 / (steps_per_real * ms_per_step * 1e-3),
        "roofline": {"bound": "hbm", "achieved": round(kern_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(kern_gbs / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": kernel, "kernel_ms": round(kernel_ms, 5),
                     "steps_per_launch": round(steps_per_launch, 3),
                     "algorithmic_bytes_per_launch": round(steps_per_launch * nv * bpv), "launch": launch},
        "cpu_baseline": cpu,
        "total_calcium": tot,
    }
    out["roofline"]["traffic_source"] = ("rocprofv3 FETCH_SIZE/WRITE_SIZE of this launch shape and kernel source "
                                         "(profiles/traffic.json)" if traffic else "no profile of this kernel source")
    if traffic:
        out["roofline"]["hbm_frac"] = round(traffic / (kernel_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
    if not args.no_extra:
        out["fp64"] = fp64_record(args.size, 20)
        out["fp64"]["fp32_error"] = fp32_error_record(args.size)
        out["g1"] = g1_record()
        out["ensemble"] = ensemble_record()
        out["closure"] = closure_record(ks=not args.no_ks, ks_count=args.ks_count)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
