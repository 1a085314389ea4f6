"""Ablation of the initial-pose stage: what the learned proposer gains over
the reference's own method.

Three proposers feed the same Anderson-accelerated ICP refine (r_max 0.05 m,
<= 50 iterations, epsilon 1e-8) of every pair, on two workloads of
asymmetric box-composite pairs (1 mm noise):
  * so3: the source turned by U[0, 180 deg] about a uniform axis (config 4);
  * yaw: the source turned about the vertical axis only (the gravity-aligned
    setting the reference's yaw search assumes).
Proposers: the learned network with 8 orientation hypotheses; the keypoint +
TDF + yaw-search coarse proposer (rtr_coarse_*, the reference's method);
and none (identity rotation about the centroids: the method without any
proposer).  A registration counts when its pose is within 1 mm RMSE of the
reference pose (ICP started at the generating pose).

    python -m realtime_robot_amd.ablation --pairs 128      # one JSON line
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import learned, synth
from .coarse import CoarseProposer
from .registration import BatchRegistration


def rotation_error_deg(T, G):
    c = (np.einsum("nij,nij->n", np.asarray(T)[:, :3, :3], np.asarray(G)[:, :3, :3]) - 1.0) / 2.0
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def run(pairs: int = 128, seed_offset: int = 0, weights=None, verbose: bool = True):
    if weights is None:
        weights, _ = learned.load_weights()
    params = learned.default_params()
    refine = learned.L.Params.from_buffer_copy(params)
    refine.anderson_memory = 5
    out = {"pairs": pairs}
    for name, gen in (("so3", synth.config4_batch), ("yaw", synth.config4_yaw_batch)):
        srcs, tgts, gts = gen(pairs, seed_offset=seed_offset)
        n = pairs
        reg = BatchRegistration()
        reg.set_clouds(srcs, tgts)
        reg.set_problems(range(n), range(n))
        reg.set_params(params)
        T_ref = [r["T"] for r in reg.align(gts)]
        reg.set_params(refine)
        cs = np.stack([s.astype(np.float64).mean(0) for s in srcs])
        ct = np.stack([t.astype(np.float64).mean(0) for t in tgts])
        props, ms = {}, {}
        lr = learned.LearnedRegistration(weights, params, max_pairs=n)
        lr.set_clouds(srcs, tgts, cs, ct)
        lr.prop.propose_multi(lr.rots, lr.tau)
        _, _, props["learned"] = lr.prop.propose_multi(lr.rots, lr.tau)
        ms["learned"] = lr.prop.last_ms()
        cp = CoarseProposer()
        cp.set_clouds(srcs, tgts)
        cp.propose()
        props["coarse"], _ = cp.propose()
        ms["coarse"] = cp.last_ms()
        none = np.tile(np.eye(4), (n, 1, 1))
        none[:, :3, 3] = ct - cs
        props["none"], ms["none"] = none, 0.0
        res = {}
        for k, P in props.items():
            reg.set_profiling(True)
            r = reg.align(P)
            t_ms = reg.timing()["ms_icp"]
            reg.set_profiling(False)
            ok = sum(synth.rmse_between(r[i]["T"], T_ref[i], srcs[i]) <= 1e-3 for i in range(n))
            rot = rotation_error_deg(P, np.stack(gts))
            res[k] = {"within_1mm": int(ok), "frac": ok / n, "proposal_rot_err_deg_median": float(np.median(rot)),
                      "proposal_within_15deg": float(np.mean(rot < 15.0)), "propose_ms": float(ms[k]),
                      "refine_ms": float(t_ms), "refine_iterations": int(sum(x["iterations"] for x in r))}
            if verbose:
                print(f"[ablation] {name} {k}: {ok}/{n} within 1 mm; proposal rotation error median "
                      f"{np.median(rot):.1f} deg; propose {ms[k]:.2f} ms, refine {t_ms:.2f} ms", file=sys.stderr,
                      flush=True)
        out[name] = res
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=128)
    ap.add_argument("--seed-offset", type=int, default=0)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    out = run(args.pairs, args.seed_offset)
    out["seconds"] = time.perf_counter() - t0
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
