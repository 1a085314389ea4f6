"""Plain GPU ICP against the FP64 oracle on the learned path's own workload.

Pairs of asymmetric box composites (source 2k-50k, target 10k-200k points,
rotation U[0, 180 deg]) are registered from the pose network's proposals,
the inits the refine starts from, by plain point-to-point ICP (r_max 0.05 m,
<= 50 iterations, eps 1e-8), and compared with the FP64 oracle's ICP
(oracle/oracle.c orc_icp, the restated pcl::IterativeClosestPoint of
RealTimeRobot/function.h:112-119) from the same proposals.  The bar is the
north star's: every transform within 1e-5 rotation (Frobenius) and 1e-5 m
translation, the same iteration count and stop rule, at the refine launch's
tuning and at tunings that change the work-item size, the search variant
(compacted walk, 8-deep gathers) and the LDS candidate tiles, since the order
in which the moments are added follows the tuning.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import realtime_robot_amd as R
from helpers import rot_err, trans_err
from realtime_robot_amd import learned, synth

pytestmark = pytest.mark.gpu

TOL_R = 1e-5
TOL_T = 1e-5
N_PAIRS = 16

# (cell_factor, cells_per_point, item_pts, blocks_per_cu, use_tile, retire, fine_from)
TUNINGS = {
    "refine": (0.25, 2.0, 1024, 0, 0, 0.5, 0),    # compacted walk, 1024-point items
    "fine": (0.25, 2.0, 1024, 0, 0, 0.5, 4),      # the same, 256-point items from iteration 4 on
    "deep": (0.35, 2.0, 512, 2, 0, 0.0, 0),       # 8-deep gathers, 2 workgroups per CU
    "tiles": (0.25, 2.0, 256, 0, 1, 0.0, 0),      # LDS candidate tiles, small items
}


@pytest.fixture(scope="module")
def workload():
    srcs, tgts, gts = synth.config4_batch(N_PAIRS)
    lr = learned.LearnedRegistration(max_pairs=N_PAIRS)
    lr.set_clouds(srcs, tgts)
    _, _, props = lr.prop.propose_multi(lr.rots, lr.tau)
    del lr
    params = learned.default_params()
    p = oracle.params(max_iter=params.max_iterations, r_max=params.max_correspondence_distance,
                      eps_T=params.transformation_epsilon)
    with ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(lambda i: oracle.icp(srcs[i], tgts[i], p, guess=props[i]), range(N_PAIRS)))
    print(f"workload: {N_PAIRS} pairs, FP64 iterations {[r['iterations'] for r in ref]}", flush=True)
    return srcs, tgts, props, ref, params


@pytest.mark.parametrize("name", list(TUNINGS))
def test_plain_icp_from_proposals_matches_fp64(workload, name):
    srcs, tgts, props, ref, params = workload
    reg = R.BatchRegistration()
    reg.set_params(params)
    reg.set_clouds(srcs, tgts)
    reg.set_problems(range(N_PAIRS), range(N_PAIRS))
    cf, cpp, ip, bpc, tile, retire, fine = TUNINGS[name]
    reg.set_tuning(cell_factor=cf, cells_per_point=cpp, item_pts=ip, blocks_per_cu=bpc, use_tile=tile,
                   retire=retire, fine_from=fine)
    assert reg.tuning()[6] == fine
    res = reg.align(props)
    again = reg.align(props)
    for i in range(N_PAIRS):
        assert np.array_equal(res[i]["T"], again[i]["T"]) and res[i]["iterations"] == again[i]["iterations"]
    worst = []
    for i in range(N_PAIRS):
        er, et = rot_err(res[i]["T"], ref[i]["T"]), trans_err(res[i]["T"], ref[i]["T"])
        worst.append((max(er, et), i, er, et, res[i]["iterations"], ref[i]["iterations"]))
    worst.sort(reverse=True)
    print(f"{name}: worst pairs (err, pair, rot, trans, iterations gpu/fp64): {worst[:3]}", flush=True)
    for _, i, er, et, it_g, it_r in worst:
        assert er <= TOL_R and et <= TOL_T, f"pair {i}: rotation {er:.2e}, translation {et:.2e}"
        assert it_g == it_r, f"pair {i}: {it_g} iterations, FP64 oracle {it_r}"
        assert res[i]["state"] == ref[i]["state"], f"pair {i}: stop rule {res[i]['state']} vs {ref[i]['state']}"
        assert synth.rmse_between(res[i]["T"], ref[i]["T"], srcs[i]) <= TOL_T
