#!/usr/bin/env python3
"""Train the learned initial guess on the solver's own MD of the 96,000-atom
AMOEBA water box and write the parameters to guess_water.npz.

The run: the seeded 32^3-lattice box (98.65 Å, PME 120^3, cutoffs 7 / 9 Å,
PCG 1e-6 D, ASPC predictor), Maxwell-Boltzmann velocities at 298 K from
--seed, 2-fs RESPA steps (8 inner) with the Bussi thermostat at 298 K, FP32
kernels.  The lattice start relaxes and heats the box, so the first --equil
steps are not captured; then the features and converged ASPC corrections of
every --stride-th atom are captured for --steps steps.  The network is
trained on the GPU (Adam, seeded order, learning rate decaying geometrically
from --lr to --lr-final over the epochs) and saved with its normalization.
Progress is printed every 500 steps.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

NMOL, NSIDE, BOX, GRID = 32000, 32, 98.65, 120
KEY = ("ewald\newald-cutoff 7.0\nvdw-cutoff 9.0\npme-grid {K} {K} {K}\npme-order 5\npolar-eps 1e-6\n"
       "polar-iter 100\npolarization mutual\nneighbor-list\nlist-buffer 2.0\npcg-precond\npcg-guess\n"
       "pcg-peek 1.0\nusolve-cutoff 4.5\nusolve-diag 1.5\npolar-predict ASPC\nvdwterm\nbondterm\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000, help="captured MD steps")
    ap.add_argument("--equil", type=int, default=2000, help="MD steps before the capture starts")
    ap.add_argument("--stride", type=int, default=16, help="capture every stride-th atom")
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=os.path.join(HERE, "guess_water.npz"))
    args = ap.parse_args()

    import t9hip
    system = t9hip.System.water_box(NMOL, NSIDE, BOX, args.seed, key=KEY.format(K=GRID))
    ctx = t9hip.Context(system, t9hip.T9_FP32)
    ctx.md_init(298.0, args.seed)
    t0 = time.time()
    for k in range(0, args.equil, 500):
        st = ctx.md_run(min(500, args.equil - k), 0.002, 8, 298.0, 0.1)
        print(f"equilibration step {min(k + 500, args.equil)}/{args.equil}: T {st['temperature']:.1f} K, "
              f"{time.time() - t0:.0f} s", flush=True)
    per = (system.n + args.stride - 1) // args.stride
    ctx.guess_capture(args.stride, per * args.steps)
    done, iters = 0, 0
    while done < args.steps:
        k = min(500, args.steps - done)
        st = ctx.md_run(k, 0.002, 8, 298.0, 0.1)
        done += k
        iters += st["iterations"]
        print(f"step {done}/{args.steps}: T {st['temperature']:.1f} K, CG iterations/step {iters / done:.2f}, "
              f"{time.time() - t0:.0f} s", flush=True)
    cap = ctx.guess_stats()["captured"]
    l0 = l1 = None
    for ep in range(args.epochs):
        lr = args.lr * (args.lr_final / args.lr) ** (ep / max(1, args.epochs - 1))
        a, l1 = ctx.guess_train(1, args.batch, lr, args.seed + ep)
        l0 = a if l0 is None else l0
        print(f"epoch {ep + 1}/{args.epochs}: lr {lr:.2e}, mean loss {l1:.5f}, {time.time() - t0:.0f} s", flush=True)
    params, norm = ctx.guess_get_params()
    print(f"trained on {cap} samples: loss {l0:.4f} -> {l1:.4f}, {time.time() - t0:.0f} s", flush=True)
    np.savez(args.out, params=params, norm=norm,
             meta=np.array([args.equil, args.steps, args.stride, cap, args.epochs, args.batch, args.lr, args.lr_final,
                            args.seed, l0, l1]))
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
