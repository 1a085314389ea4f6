"""What the judge accepts and the 1 mm bar rejects, on the bench's workload.

256 config-4 pairs through the learned pipeline (proposal, Anderson +
coarse-to-fine refine, judge, fallback).  A registration misses when it lies
more than 1 mm RMSE (over its source points) from the FP64 oracle's ICP
started at the generating pose (oracle/oracle.c orc_icp, the restated
pcl::IterativeClosestPoint of RealTimeRobot/function.h:112-119).  For every
miss this test checks that the difference is a slide along the surfaces,
which no residual-based acceptance test can see:

  * the displacement of the source points between the two poses, projected
    on the target's normals at the reference correspondences, is at most
    half of its RMS (measured on the 20 misses: 26-45 %, 0.33-0.71 mm
    against 1.01-1.96 mm);
  * point-to-point ICP's own objective (mean squared distance to the nearest
    target point within r_max) is no more than 2 % higher at the learned
    pose than at the reference pose (measured: at most 1.4 % higher, and
    lower for 2 of the 20);
  * at most 32 of 256 registrations miss (measured 20).

The objective is flat to within about 1 % along 1-2 mm tangential slides of
these box composites (noise 1 mm), and ICP's stop rule (transformation
epsilon) ends the slide wherever the increments become small, so two ICP
runs from different starts stop at different points of the same valley.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from realtime_robot_amd import learned, synth

pytestmark = pytest.mark.gpu

N = 256
R_MAX = 0.05


def _mse(src, kd, T):
    q = (src.astype(np.float64) @ T[:3, :3].T + T[:3, 3]).astype(np.float32)
    idx, d2 = kd.nn(q, r_max=R_MAX)
    ok = idx >= 0
    return float(np.mean(d2[ok].astype(np.float64))), ok


def test_accepted_misses_are_tangential_slides():
    srcs, tgts, gts = synth.config4_batch(N)
    cs = np.stack([s.astype(np.float64).mean(0) for s in srcs])
    ct = np.stack([t.astype(np.float64).mean(0) for t in tgts])
    lr = learned.LearnedRegistration(max_pairs=N)
    lr.set_clouds(srcs, tgts, cs, ct)
    res = lr.align()
    T = [r["T"] for r in res]
    p = oracle.params(max_iter=50, r_max=R_MAX, eps_T=1e-8)
    with ThreadPoolExecutor(16) as ex:
        ref = list(ex.map(lambda i: oracle.icp(srcs[i], tgts[i], p, guess=gts[i])["T"], range(N)))
    err = np.array([synth.rmse_between(T[i], ref[i], srcs[i]) for i in range(N)])
    miss = np.nonzero(err > 1e-3)[0]
    print(f"{len(miss)} of {N} registrations beyond 1 mm of the FP64 registration from the generating pose "
          f"({lr.stats['n_fallback']} fallbacks)", flush=True)
    assert len(miss) <= 32

    def check(i):
        kd = oracle.KdTree(tgts[i])
        s = srcs[i].astype(np.float64)
        q_ref = s @ ref[i][:3, :3].T + ref[i][:3, 3]
        q_l = s @ T[i][:3, :3].T + T[i][:3, 3]
        idx, _ = kd.nn(q_ref.astype(np.float32), r_max=R_MAX)
        ok = idx >= 0
        nrm, _ = oracle.normals(tgts[i], 20)
        d = (q_l - q_ref)[ok]
        total = np.sqrt(np.mean(np.sum(d * d, 1)))
        normal = np.sqrt(np.mean(np.sum(d * nrm[idx[ok]].astype(np.float64), 1) ** 2))
        mse_l, _ = _mse(srcs[i], kd, T[i])
        mse_r, _ = _mse(srcs[i], kd, ref[i])
        return i, err[i], total, normal, mse_l / mse_r

    with ThreadPoolExecutor(8) as ex:
        rows = list(ex.map(check, miss))
    for i, e, total, normal, rel in rows:
        print(f"pair {i}: {e * 1e3:.2f} mm from the reference, normal component {normal * 1e3:.2f} mm "
              f"({normal / total:.2f}), ICP objective x{rel:.4f}", flush=True)
        assert normal <= 0.5 * total, f"pair {i}: not a tangential slide"
        assert rel <= 1.02, f"pair {i}: the objective is {rel:.3f}x the reference's"
