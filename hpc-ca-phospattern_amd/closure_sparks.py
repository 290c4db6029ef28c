#!/usr/bin/env python3
"""Spark statistics of the closure form, checked against the plain fine-grid method.

The same seeded dyads (12 nm geometries from dyad.ensemble_dyad) run three
ways, each as one ensemble:

  fine     plain method at 3 nm: the reference refines the 12 nm geometry
           (species.c:938-1035) and gates each channel on its own cleft
           voxel (boundaryfluxes.c:295-359);
  closure  12 nm, every channel gated on the closure network's prediction
           of the fine mouth Ca (Ensemble.GATE_CLOSURE) and its flux scaled
           by the flux network's fine-to-coarse release ratio
           (Ensemble.set_flux_closure);
  gating   the same without the flux network (gating closure only);
  coarse   plain method at 12 nm (gating on the coarse cleft voxel).

Every realisation gives peak dF/F0, duration (FDHM) and RyRs recruited
(sparks.py) from a linescan of the dye through the cluster centre along x,
sampled at the same times on both grids (one coarse step is 16 fine
steps).  The closure, gating and coarse distributions are compared with
the fine one by two-sample Kolmogorov-Smirnov tests.  Realisations whose closure
inputs left the training range at any evaluation (the range monitor) are
run again with the fine method, and their statistics replace the closure
ones: the fine-grid fallback.

The gating is the fast kinetics of dyad.fast_spark_dyad (opening at up to
200/ms above a few uM, closing at 50/ms), with the phosphorylated RyRs
twice as sensitive, so a spark runs its course in a few tenths of a
millisecond and the 3 nm runs stay affordable.

    python closure_sparks.py [--count N] [--edge LO HI] [--ryr LO HI] [--tstop MS] [--json OUT]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import closure  # noqa: E402
import csparks as cs  # noqa: E402
import dyad  # noqa: E402
import sparks  # noqa: E402

FAST = dict(Kd_open=5.0, k_max_open=200.0, k_min_close=50.0, k_max_close=50.0)
DYE = 1                  # linescan species (Fluo)
FINE_H, COARSE_H = 3.0, 12.0
KEYS = ("peak", "duration", "recruited")


def spark_model(model: cs.SpeciesModel) -> cs.SpeciesModel:
    """The dyad model with the fast spark kinetics (pRyR at half the Kd)."""
    ryr = dataclasses.replace(model.ryr, **FAST)
    pryr = dataclasses.replace(model.p_ryr, **{**FAST, "Kd_open": FAST["Kd_open"] / 2}) if model.p_ryr else None
    return dataclasses.replace(model, ryr=ryr, p_ryr=pryr)


def spark_dyads(seeds: Sequence[int], edge: Tuple[float, float], ryr: Tuple[int, int]) -> List[dyad.Dyad]:
    out = []
    for s in seeds:
        d = dyad.ensemble_dyad(s, h=COARSE_H, edge_nm=edge, ryr=ryr)
        out.append(dyad.Dyad(d.geometry, spark_model(d.model), d.open_index, d.description))
    return out


def line_offsets(d: dyad.Dyad, h: float) -> Tuple[int, int, int]:
    """Linescan offsets (from the grid centre, in voxels of size h) of the
    line along x through the cluster centre in the middle of the cleft.
    Every ensemble dyad has its cluster at the centre of the xy plane and
    the same layers along z, so the offsets are the same for all."""
    g = d.geometry
    r = int(round(g.h / h))
    cleft = np.argwhere(g.voxels == dyad.CLEFT)
    zc = (int(cleft[:, 2].min()) + int(cleft[:, 2].max()) + 1) * r // 2   # middle fine plane of the cleft
    nz = g.voxels.shape[2] * r
    return (0, 0, zc - nz // 2)


@dataclass
class Run:
    stats: List[sparks.SparkStats]
    flagged: np.ndarray          # per realisation: closure evaluations out of range (0 without closure)
    voxels: int
    steps: int
    seconds: float


def run(dyads: Sequence[dyad.Dyad], seeds: Sequence[int], h: float, nsteps: int, save_every: int,
        gating: int = cs.Ensemble.GATE_OWN, net: Optional[cs.Mlp] = None, precision: str = "f32",
        graphs: bool = True, say=None, flux_net: Optional[cs.Mlp] = None) -> Run:
    """Spark statistics of every dyad at voxel size h, one ensemble (steps
    replayed as captured graphs when `graphs`)."""
    ens = cs.Ensemble([(d.geometry, d.model, cs.Options(h=h, seed=s, open={"ryr": [d.open_index]},
                                                        precision=precision))
                       for d, s in zip(dyads, seeds)], instances=list(range(len(dyads))))
    if gating != cs.Ensemble.GATE_OWN:
        ens.set_gating(gating, net)
    if flux_net is not None:
        ens.set_flux_closure(flux_net)
    best, _ = ens.autotune(2)            # the stencil pass's launch shape for these grids
    ens.set_launch(**best)
    if graphs:
        ens.set_graph(True)
    _, nv, _ = ens.info()
    t0 = time.perf_counter()
    last = [t0]

    def progress(done, total):
        now = time.perf_counter()
        if say is not None and now - last[0] > 20.0:
            say(f"  h={h:g} nm: {done}/{total} steps, {now - t0:.0f} s")
            last[0] = now

    stats = sparks.realise_ensemble(ens, nsteps, save_every, DYE, line_offsets(dyads[0], h), dyads[0].model.conv,
                                    progress=progress)
    ens.synchronize()
    sec = time.perf_counter() - t0
    flagged = ens.monitor() if gating == cs.Ensemble.GATE_CLOSURE else np.zeros(len(dyads), np.uint32)
    ens.close()
    return Run(stats, flagged, nv, nsteps, sec)


def ks(a: Sequence[sparks.SparkStats], b: Sequence[sparks.SparkStats]) -> dict:
    from scipy.stats import ks_2samp
    out = {}
    for k in KEYS:
        x = np.array([getattr(s, k) for s in a], dtype=float)
        y = np.array([getattr(s, k) for s in b], dtype=float)
        r = ks_2samp(x, y)
        out[k] = {"D": float(r.statistic), "p": float(r.pvalue), "mean": float(x.mean()), "ref_mean": float(y.mean())}
    return out


def compare(count: int = 256, edge=(192.0, 256.0), ryr=(9, 30), tstop_ms: float = 0.3, save_coarse: int = 20,
            seed0: int = 1, net: Optional[cs.Mlp] = None, flux_net: Optional[cs.Mlp] = None,
            closure_graphs: bool = True, say=print) -> dict:
    """Fine, closure (with fallback), gating-only and coarse statistics of
    `count` realisations, their KS comparisons and throughputs.  Without
    `net`, both networks come from closure_params.npz."""
    seeds = list(range(seed0, seed0 + count))
    dyads = spark_dyads(seeds, edge, ryr)
    own_net = net is None
    fnet = flux_net
    if own_net:
        path = os.path.join(HERE, "closure_params.npz")
        net, _, trained = closure.load(path)
        if not trained:
            raise RuntimeError("closure_params.npz is missing: train the closure first (train_closure.py)")
        fl = closure.load_flux(path)
        fnet = fl[0] if fl is not None else None
    probe_f = cs.Simulation(dyads[0].geometry, dyads[0].model, cs.Options(h=FINE_H, seed=1))
    probe_c = cs.Simulation(dyads[0].geometry, dyads[0].model, cs.Options(h=COARSE_H, seed=1))
    ratio = probe_c.dt / probe_f.dt
    dt_c = probe_c.dt
    probe_f.close()
    probe_c.close()
    r = int(round(ratio))
    if abs(ratio - r) > 1e-9 * ratio:
        raise ValueError(f"coarse dt is not a whole number of fine steps ({ratio})")
    n_coarse = int(np.ceil(tstop_ms / dt_c / save_coarse)) * save_coarse
    say(f"{count} dyads (edges {edge[0]:.0f}-{edge[1]:.0f} nm, {ryr[0]}-{ryr[1]} RyRs), {n_coarse} coarse / "
        f"{n_coarse * r} fine steps ({n_coarse * dt_c:.4f} ms), linescan every {save_coarse} coarse steps")
    res = {}
    res["coarse"] = run(dyads, seeds, COARSE_H, n_coarse, save_coarse, say=say)
    say(f"coarse: {res['coarse'].seconds:.1f} s")
    res["gating"] = run(dyads, seeds, COARSE_H, n_coarse, save_coarse, cs.Ensemble.GATE_CLOSURE, net,
                        graphs=closure_graphs, say=say)
    say(f"gating closure only: {res['gating'].seconds:.1f} s")
    if fnet is not None:
        res["closure"] = run(dyads, seeds, COARSE_H, n_coarse, save_coarse, cs.Ensemble.GATE_CLOSURE, net,
                             graphs=closure_graphs, say=say, flux_net=fnet)
    else:
        say("no flux network in closure_params.npz: the closure is the gating closure alone")
        res["closure"] = res["gating"]
    say(f"closure: {res['closure'].seconds:.1f} s, {int(np.count_nonzero(res['closure'].flagged))} realisations "
        "flagged by the range monitor")
    res["fine"] = run(dyads, seeds, FINE_H, n_coarse * r, save_coarse * r, say=say)
    say(f"fine: {res['fine'].seconds:.1f} s")
    # fallback: flagged realisations take the fine method's statistics
    flagged = np.flatnonzero(res["closure"].flagged)
    closure_stats = list(res["closure"].stats)
    fallback_s = 0.0
    if len(flagged):
        fb = run([dyads[i] for i in flagged], [seeds[i] for i in flagged], FINE_H, n_coarse * r, save_coarse * r,
                 say=say)
        fallback_s = fb.seconds
        for k, i in enumerate(flagged):
            closure_stats[i] = fb.stats[k]
    if own_net:
        net.close()
        if fnet is not None:
            fnet.close()
    out = {"count": count, "edge_nm": list(edge), "ryr": list(ryr), "tstop_ms": n_coarse * dt_c,
           "coarse_steps": n_coarse, "fine_steps": n_coarse * r, "kinetics": FAST,
           "flagged": int(len(flagged)), "flux_closure": fnet is not None}
    for name in ("fine", "closure", "gating", "coarse"):
        rr = res[name]
        sec = rr.seconds + (fallback_s if name == "closure" else 0.0)
        out[name] = {"voxels": rr.voxels, "seconds": round(sec, 3), "realisations_per_s": round(count / sec, 3)}
    out["closure"]["ks_vs_fine"] = ks(closure_stats, res["fine"].stats)
    out["gating"]["ks_vs_fine"] = ks(res["gating"].stats, res["fine"].stats)
    out["coarse"]["ks_vs_fine"] = ks(res["coarse"].stats, res["fine"].stats)
    out["speedup_closure_vs_fine"] = round(out["closure"]["realisations_per_s"] / out["fine"]["realisations_per_s"], 1)
    for name in ("closure", "gating", "coarse"):
        for k in KEYS:
            q = out[name]["ks_vs_fine"][k]
            say(f"{name:8s} {k:9s}: mean {q['mean']:.4g} vs fine {q['ref_mean']:.4g}, KS D {q['D']:.4f} p {q['p']:.4f}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--edge", type=float, nargs=2, default=(192.0, 256.0))
    ap.add_argument("--ryr", type=int, nargs=2, default=(9, 30))
    ap.add_argument("--tstop", type=float, default=0.3)
    ap.add_argument("--seed0", type=int, default=1)
    ap.add_argument("--json", default="")
    ap.add_argument("--no-closure-graphs", action="store_true", help="run the closure ensemble step by step")
    args = ap.parse_args()
    out = compare(args.count, tuple(args.edge), tuple(args.ryr), args.tstop, seed0=args.seed0,
                  closure_graphs=not args.no_closure_graphs, say=lambda m: print(m, flush=True))
    print(json.dumps(out), flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
