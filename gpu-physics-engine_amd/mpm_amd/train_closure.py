"""Train the learned closure for the learned1024 workload and write its parameters.

    python gpu-physics-engine_amd/mpm_amd/train_closure.py --out gpu-physics-engine_amd/mpm_amd/data/closure_learned1024.npy

Pairs: the states of every jelly / snow particle of the 4,194,304-particle jelly + snow scene after each of
`--ticks` ticks of the plain (analytic) solver, plus random deformation gradients (singular values log-uniform
in [0.6, 1.6]) up to `--pairs` in total; targets are the analytic tau / mu.  Training: `--steps` Adam
minibatches on the GPU (mpm_closure_train).  Writes the float32[4675] parameters and a JSON report next to them.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from mpm_amd import abi, closure, scenes  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "data", "closure_learned1024.npy"))
    ap.add_argument("--ticks", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=3)
    args = ap.parse_args()
    t0 = time.time()
    p = abi.preset(abi.PRESET_LEARNED1024)
    sc = scenes.for_preset(abi.PRESET_LEARNED1024)
    harvested = closure.harvest(p, sc, ticks=args.ticks)
    n_h = harvested[0].shape[0]
    X, T = closure.dataset(p, harvested, n_random=max(0, args.pairs - n_h), seed=args.seed + 100)
    print(f"[train_closure] {X.shape[0]} in-range pairs ({n_h} harvested states) in {time.time() - t0:.1f} s",
          flush=True)
    lr_sched = [(args.steps // 2, args.lr), (args.steps // 4, args.lr / 4), (args.steps - args.steps // 2 - args.steps // 4,
                                                                             args.lr / 16)]
    w = None
    losses = []
    for k, (steps, lr) in enumerate(lr_sched):
        w, loss = closure.train(X, T, steps=steps, batch=args.batch, lr=lr, seed=args.seed + k, weights=w)
        losses.append(loss)
        print(f"[train_closure] phase {k}: {steps} steps at lr {lr:g}, minibatch loss {loss:.3e}", flush=True)
    # held-out random deformations
    F, J, M = closure.random_deformations(200000, seed=args.seed + 999)
    Xh, Th, V = closure.samples(p, F, J, M)
    Xh, Th = Xh[V], Th[V]
    err = np.sqrt(np.mean((closure.evaluate(Xh, w) - Th) ** 2))
    scale = np.sqrt(np.mean(Th ** 2))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, w.astype(np.float32))
    report = {"pairs": int(X.shape[0]), "harvested_states": int(n_h), "steps": args.steps, "batch": args.batch,
              "losses": losses, "heldout_rms_error": float(err), "heldout_rms_target": float(scale),
              "seconds": time.time() - t0}
    json.dump(report, open(os.path.splitext(args.out)[0] + ".json", "w"), indent=1)
    print(f"[train_closure] held-out rms error {err:.3e} (target rms {scale:.3e}); wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
