"""Train the learned pose proposer on the device.

    python -m realtime_robot_amd.train_proposer --base 12500 --steps 8000 --out realtime_robot_amd/weights/proposer.npz

Builds the labelled base set (canonical box-composite pairs, labels = the
GPU ICP's converged pose, device farthest-point sampling), trains with Adam
on batches of 256 augmented pairs (every pair's source moved by a fresh
rigid motion with rotation U[0, 180 deg]), reports the proposal error on
held-out pairs and writes the FP32 weights with their training record.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--base", type=int, default=12500, help="labelled base pairs")
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--holdout", type=int, default=512, help="held-out base pairs (separate seed)")
    ap.add_argument("--out", type=str, required=True)
    args = ap.parse_args(argv)

    from realtime_robot_amd import learned

    def log(msg):
        print(f"[train] {msg}", flush=True)

    t0 = time.time()
    ts = learned.build_training_set(args.base, seed=20240 + args.seed, log=log)
    ho = learned.build_training_set(args.holdout, seed=777777, log=log)
    log(f"base set {len(ts)} pairs, holdout {len(ho)} ({time.time() - t0:.1f} s)")
    w, hist = learned.train(ts, args.steps, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    err = learned.evaluate(w, ho, n=2048)
    qs = np.percentile(err, [10, 25, 50, 75, 90])
    log(f"holdout rotation error deg: p10 {qs[0]:.1f} p25 {qs[1]:.1f} median {qs[2]:.1f} p75 {qs[3]:.1f} "
        f"p90 {qs[4]:.1f}; <15deg {np.mean(err < 15):.3f} <30deg {np.mean(err < 30):.3f}")
    meta = {"base_pairs": len(ts), "steps": args.steps, "batch": args.batch, "lr": args.lr, "seed": args.seed,
            "pairs_seen": args.steps * args.batch, "final_loss": float(np.mean(hist[-200:])),
            "holdout_rot_err_deg_percentiles_10_25_50_75_90": [float(v) for v in qs],
            "holdout_frac_below_15deg": float(np.mean(err < 15)), "holdout_frac_below_30deg": float(np.mean(err < 30)),
            "train_seconds": time.time() - t0,
            "loss_curve": [float(np.mean(hist[i:i + 100])) for i in range(0, len(hist), max(1, len(hist) // 40))]}
    learned.save_weights(args.out, w, meta)
    log(f"wrote {args.out} ({time.time() - t0:.1f} s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
