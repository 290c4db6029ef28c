#!/usr/bin/env python3
"""Train the closure networks on paired 3 nm / 12 nm runs of generated
12 nm dyads and write their parameters (closure.save): the gating network
(fine Ca at each channel mouth) and the flux network (fine-to-coarse
release ratio of each open channel).

    python train_closure.py [out.npz] [--dyads N] [--coarse-steps K] [--steps S] [--edge LO HI]
                            [--ryr LO HI] [--sample-every K]

The fine ensemble gates with Ca-independent rates (closure.schedule_model)
so that the samples cover many open/closed configurations; the coarse
ensemble replays its channel states.  Held-out dyads give the errors of
both networks and of the plain coarse grid, printed at the end.

The defaults are the closure-training workload: 2048 dyads of the
ensemble generator (edges 256-512 nm, 9-100 RyRs), 300 coarse steps,
sampled every 16: about 2.1 million samples, 4 minutes on one MI355X."""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import closure  # noqa: E402
import dyad  # noqa: E402
import mlp_init  # noqa: E402


def dyads(seeds, edge, ryr):
    return [dyad.ensemble_dyad(s, h=12.0, edge_nm=edge, ryr=ryr) for s in seeds]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out", nargs="?", default=os.path.join(HERE, "closure_params.npz"))
    ap.add_argument("--dyads", type=int, default=2048)
    ap.add_argument("--coarse-steps", type=int, default=300)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--edge", type=float, nargs=2, default=(256.0, 512.0))
    ap.add_argument("--ryr", type=int, nargs=2, default=(9, 100))
    ap.add_argument("--sample-every", type=int, default=16, help="coarse steps between samples")
    args = ap.parse_args()
    say = lambda m: print(m, flush=True)
    t0 = time.time()
    train = closure.paired_samples(dyads(range(args.dyads), args.edge, args.ryr), list(range(args.dyads)),
                                   coarse_steps=args.coarse_steps, sample_every=args.sample_every, say=say)
    print(f"{len(train.ca_fine)} training samples from {args.dyads} dyads "
          f"({train.fine_steps} fine / {train.coarse_steps} coarse steps) in {time.time() - t0:.1f}s", flush=True)
    t1 = time.time()
    net, norm = closure.train(train, mlp_init.params(0), steps=args.steps, batch=2048, lr=2e-3)
    fnet, fnorm = closure.train_flux(train, norm, mlp_init.params(1), steps=args.steps, batch=2048, lr=2e-3)
    print(f"trained 2 x {args.steps} Adam steps in {time.time() - t1:.1f}s "
          f"({int(closure.flux_selection(train).sum())} open-channel samples for the flux network)", flush=True)
    held = closure.paired_samples(dyads(range(10000, 10012), args.edge, args.ryr), list(range(10000, 10012)),
                                  coarse_steps=args.coarse_steps // 2, sample_every=4, say=say)
    e_c, e_p = closure.log_errors(net, norm, held)
    print(f"held-out log-Ca MSE at the channel mouth: closure {e_c:.5f}, plain coarse grid {e_p:.5f}", flush=True)
    f_c, f_p = closure.flux_errors(fnet, fnorm, held)
    print(f"held-out MSE of the log release ratio: flux closure {f_c:.5f}, plain coarse grid {f_p:.5f}", flush=True)
    closure.save(args.out, net, norm, flux=(fnet, fnorm), held_out_mse=np.array([e_c, e_p]),
                 flux_held_out_mse=np.array([f_c, f_p]), samples=np.array([len(train.ca_fine)]))
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
