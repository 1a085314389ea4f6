"""GPU parity tests of the registration path against the CPU oracle.

Bars: correspondence indices and squared distances bit-exact against brute
force (lowest-index ties); transforms within 1e-5 rotation (Frobenius) and
1e-5 m translation of the FP64 oracle ICP; the transformed output cloud
bit-exact against the oracle's float transform of the same T.
"""
import numpy as np
import pytest

import oracle
import realtime_robot_amd as R
from helpers import golden, rot_err, trans_err
from pcd_np import read_xyz
from realtime_robot_amd import multistart, synth

pytestmark = pytest.mark.gpu

TOL_R = 1e-5
TOL_T = 1e-5


def _icp(src, tgt, max_dist=None, iters=None, eps=None):
    icp = R.IterativeClosestPoint()
    icp.setInputSource(src)
    icp.setInputTarget(tgt)
    if max_dist is not None:
        icp.setMaxCorrespondenceDistance(max_dist)
    if iters is not None:
        icp.setMaximumIterations(iters)
    if eps is not None:
        icp.setTransformationEpsilon(eps)
    return icp


def _perturb(T, deg, t, seed):
    rng = np.random.default_rng(seed)
    D = synth.random_rigid(rng, deg, t)
    return D @ T


@pytest.mark.parametrize("max_dist", [0.05, 0.02, None])
def test_correspondences_bitexact_config1(max_dist):
    src, tgt, T_gt = synth.config1_pair(0)
    icp = _icp(src, tgt, max_dist=max_dist)
    for k, T in enumerate([np.eye(4), T_gt, _perturb(T_gt, 3.0, 0.01, 7)]):
        idx, d2 = icp.getCorrespondences(T)
        q = oracle.transform(T, src)
        i_ref, d_ref = oracle.nn_brute(q, tgt, max_dist)
        assert np.array_equal(idx, i_ref), f"T{k}: {np.sum(idx != i_ref)} index mismatches"
        assert np.array_equal(d2, d_ref), f"T{k}: squared distances differ"


def test_correspondences_ties_and_nans():
    rng = np.random.default_rng(5)
    base = np.round(rng.random((4000, 3)) * 20) / 20  # grid-aligned: many exact ties
    tgt = np.concatenate([base, base[:1000]]).astype(np.float32)  # duplicates
    tgt[17] = np.nan
    tgt[123] = [np.inf, 0, 0]
    src = (np.round(rng.random((3000, 3)) * 40) / 40).astype(np.float32)
    src[5] = np.nan
    icp = _icp(src, tgt, max_dist=0.06)
    idx, d2 = icp.getCorrespondences(np.eye(4))
    i_ref, d_ref = oracle.nn_brute(src, tgt, 0.06)
    assert np.array_equal(idx, i_ref)
    assert np.array_equal(d2, d_ref)
    assert idx[5] == -1


def test_known_answer_pure_translation():
    """5.pcd -> mcloud.pcd is a pure translation (0.8, 0.8, 0) (index-aligned
    output of an earlier reference run); plain ICP with PCL's unbounded
    correspondence distance recovers it."""
    a = read_xyz(golden("5.pcd"))
    b = read_xyz(golden("mcloud.pcd"))
    icp = _icp(a, b, iters=50)
    out = icp.align()
    T = icp.getFinalTransformation()
    assert abs(T[0, 3] - 0.8) < 1e-5 and abs(T[1, 3] - 0.8) < 1e-5 and abs(T[2, 3]) < 1e-5
    assert rot_err(T, np.eye(4)) < 1e-5
    ref = oracle.icp(a, b, oracle.params(max_iter=50))
    assert rot_err(T, ref["T"]) < TOL_R and trans_err(T, ref["T"]) < TOL_T
    assert np.array_equal(out, oracle.transform(T, a))


def test_align_config1_matches_oracle():
    src, tgt, T_gt = synth.config1_pair(0)
    icp = _icp(src, tgt, max_dist=0.05, iters=50, eps=1e-8)
    out = icp.align()
    T = icp.getFinalTransformation()
    res = icp.getResult()
    ref = oracle.icp(src, tgt, oracle.params(max_iter=50, r_max=0.05, eps_T=1e-8))
    print("gpu", res["iterations"], res["state"], "oracle", ref["iterations"], ref["state"],
          "rot", rot_err(T, ref["T"]), "t", trans_err(T, ref["T"]))
    assert rot_err(T, ref["T"]) < TOL_R and trans_err(T, ref["T"]) < TOL_T
    assert res["iterations"] == ref["iterations"]
    assert res["state"] == ref["state"]
    assert res["n_corr"] == ref["n_corr"]
    assert abs(res["mse"] - ref["mse"]) <= 1e-6 * ref["mse"]
    assert np.array_equal(out, oracle.transform(T, src))
    # ICP moved the source towards the ground truth
    assert synth.rmse_between(T, T_gt, src) < 0.5 * synth.rmse_between(np.eye(4), T_gt, src)
    f = icp.getFitnessScore()
    f_ref = oracle.fitness(src, tgt, T)
    assert abs(f - f_ref) <= 1e-9 + 1e-6 * f_ref


def test_align_with_guess_and_max_range_fitness():
    src, tgt, T_gt = synth.config1_pair(1)
    guess = _perturb(T_gt, 5.0, 0.02, 3)
    icp = _icp(src, tgt, max_dist=0.05, iters=30, eps=1e-8)
    icp.align(guess)
    T = icp.getFinalTransformation()
    ref = oracle.icp(src, tgt, oracle.params(max_iter=30, r_max=0.05, eps_T=1e-8), guess=guess)
    assert rot_err(T, ref["T"]) < TOL_R and trans_err(T, ref["T"]) < TOL_T
    for mr in (1e-6, 1e-5, 4e-6):
        f = icp.getFitnessScore(mr)
        f_ref = oracle.fitness(src, tgt, T, mr)
        assert (f == f_ref) if f_ref == np.finfo(np.float64).max else abs(f - f_ref) <= 1e-6 * f_ref


def test_no_correspondences_is_reported():
    src, tgt, _ = synth.config1_pair(2)
    icp = _icp(src + np.float32(10.0), tgt, max_dist=0.01, iters=10)
    icp.align()
    r = icp.getResult()
    assert r["state"] == R._lib.FAIL_NO_CORRESPONDENCES and not icp.hasConverged()
    assert np.allclose(icp.getFinalTransformation(), np.eye(4))
    ref = oracle.icp(src + np.float32(10.0), tgt, oracle.params(max_iter=10, r_max=0.01))
    assert ref["state"] == oracle.FAIL_NO_CORR


def test_empty_inputs_raise():
    icp = R.IterativeClosestPoint()
    icp.setInputSource(np.zeros((0, 3), np.float32))
    icp.setInputTarget(np.random.rand(100, 3).astype(np.float32))
    with pytest.raises(R.RtrError) as e:
        icp.align()
    assert e.value.code == R._lib.ERR_EMPTY_INPUT


def test_batch_matches_oracle_and_graph_replay_is_stable():
    srcs, tgts, gts = synth.config3_batch(6, seed_offset=500, ns_range=(2000, 6000), nt_range=(8000, 20000),
                                          max_deg=10.0, max_t=0.05)
    b = R.BatchRegistration()
    b.set_clouds(srcs, tgts)
    b.set_problems(np.arange(6), np.arange(6))
    b.set_params(R.make_params(max_iterations=40, max_correspondence_distance=0.05, transformation_epsilon=1e-8))
    r1 = b.align()
    r2 = b.align()
    b.set_graph(False)
    r3 = b.align()
    for p in range(6):
        for r in (r2, r3):
            assert np.array_equal(r[p]["T"], r1[p]["T"]), "replay is not deterministic"
        ref = oracle.icp(srcs[p], tgts[p], oracle.params(max_iter=40, r_max=0.05, eps_T=1e-8))
        T = r1[p]["T"]
        print(p, r1[p]["iterations"], ref["iterations"], rot_err(T, ref["T"]), trans_err(T, ref["T"]))
        assert rot_err(T, ref["T"]) < TOL_R and trans_err(T, ref["T"]) < TOL_T
        assert r1[p]["iterations"] == ref["iterations"]


def test_batch_shared_clouds_multistart():
    src, tgt, T_gt = synth.config1_pair(3)
    b = R.BatchRegistration()
    b.set_clouds([src], [tgt])
    b.set_problems([0, 0, 0], [0, 0, 0])
    b.set_params(R.make_params(max_iterations=25, max_correspondence_distance=0.05, transformation_epsilon=1e-8))
    guesses = [np.eye(4), _perturb(T_gt, 4, 0.01, 1), T_gt]
    res = b.align(guesses)
    for g, r in zip(guesses, res):
        ref = oracle.icp(src, tgt, oracle.params(max_iter=25, r_max=0.05, eps_T=1e-8), guess=g)
        assert rot_err(r["T"], ref["T"]) < TOL_R and trans_err(r["T"], ref["T"]) < TOL_T


def test_batch_resize_drops_stale_problems():
    """Replacing the clouds with fewer clouds drops a problem set that names
    clouds which no longer exist (regression: it was re-laid out with
    out-of-range cloud indices); a set that still fits is kept."""
    srcs, tgts, gts = synth.config3_batch(6, seed_offset=950, ns_range=(2000, 3000), nt_range=(4000, 6000))
    b = R.BatchRegistration()
    b.set_clouds(srcs, tgts)
    b.set_problems(range(6), range(6))
    b.set_params(R.make_params(max_iterations=10, max_correspondence_distance=0.05))
    b.align(gts)
    b.set_clouds(srcs[:2], tgts[:2])
    with pytest.raises(R.RtrError) as e:
        b.align(gts[:2])
    assert e.value.code == R._lib.ERR_NOT_READY
    b.set_problems(range(2), range(2))
    r2 = b.align(gts[:2])
    b.set_clouds(srcs[2:4], tgts[2:4])        # same count: problems 0, 1 still valid
    r3 = b.align(gts[2:4])
    assert len(r2) == len(r3) == 2
    for p in range(2):
        ref = oracle.icp(srcs[2 + p], tgts[2 + p], oracle.params(max_iter=10, r_max=0.05), guess=gts[2 + p])
        assert rot_err(r3[p]["T"], ref["T"]) < TOL_R and trans_err(r3[p]["T"], ref["T"]) < TOL_T


def test_batch_reused_grids_match_rebuild_always():
    """With the rebuild-always policy off, grids built once (build(), or the
    first align after new clouds) serve later aligns across set_problems; the
    results are bit-identical to an engine that rebuilds on every align."""
    srcs, tgts, gts = synth.config3_batch(4, seed_offset=960, ns_range=(2000, 4000), nt_range=(5000, 9000))
    p = R.make_params(max_iterations=15, max_correspondence_distance=0.05, transformation_epsilon=1e-8)
    a, b = R.BatchRegistration(), R.BatchRegistration()
    for e in (a, b):
        e.set_clouds(srcs, tgts)
        e.set_params(p)
    b.set_rebuild(False)
    b.build()
    for ids, guesses in [([0, 1, 2, 3], gts), ([2, 2, 3], [gts[2], np.eye(4), gts[3]]), ([1], [gts[1]])]:
        for e in (a, b):
            e.set_problems(ids, ids)
        ra, rb = a.align(guesses), b.align(guesses)
        for x, y in zip(ra, rb):
            assert np.array_equal(x["T"], y["T"]) and x["iterations"] == y["iterations"]
    # new clouds: the next align rebuilds by itself
    b.set_clouds(srcs[::-1], tgts[::-1])
    b.set_problems([0], [0])
    r = b.align([gts[3]])[0]
    ref = oracle.icp(srcs[3], tgts[3], oracle.params(max_iter=15, r_max=0.05, eps_T=1e-8), guess=gts[3])
    assert rot_err(r["T"], ref["T"]) < TOL_R and trans_err(r["T"], ref["T"]) < TOL_T


def test_anderson_reaches_an_equal_minimum_in_fewer_iterations():
    """Anderson-accelerated ICP (anderson_memory > 0, no PCL counterpart)
    reaches a minimum as good as the plain iteration's.  Run both to a
    vanishing transformation epsilon from 5-degree perturbations: the final
    MSE of the accelerated run is within 0.1% of the plain one's and its pose
    within 1 mm RMSE of the plain fixed point.  (The MSE of these box
    composites, with float correspondences, is flat to ~1e-9 m^2 over a few
    tenths of a millimetre along sliding directions and holds shallow local
    minima there; the two paths can settle in different ones.)  At
    the workload's epsilon (1e-8) the accelerated runs take under 80% of the
    plain iterations in total, stop no farther from the fixed point on average,
    replay bit-identically from the graph, and invalid memories are refused."""
    srcs, tgts, gts = synth.config3_batch(8, seed_offset=970, ns_range=(3000, 8000), nt_range=(10000, 30000),
                                          max_deg=10.0, max_t=0.05)
    guesses = [_perturb(gts[i], 5.0, 0.01, 100 + i) for i in range(8)]
    b = R.BatchRegistration()
    b.set_clouds(srcs, tgts)
    b.set_problems(range(8), range(8))
    tight = dict(max_iterations=3000, max_correspondence_distance=0.05, transformation_epsilon=1e-18)
    b.set_params(R.make_params(**tight))
    plain = b.align(guesses)
    b.set_params(R.make_params(anderson_memory=5, **tight))
    acc = b.align(guesses)
    for p in range(8):
        d = synth.rmse_between(plain[p]["T"], acc[p]["T"], srcs[p])
        print(p, plain[p]["iterations"], acc[p]["iterations"], plain[p]["mse"], acc[p]["mse"], d)
        assert acc[p]["mse"] <= plain[p]["mse"] * (1 + 1e-3)
        assert d < 1e-3
    work = dict(max_iterations=50, max_correspondence_distance=0.05, transformation_epsilon=1e-8)
    b.set_params(R.make_params(**work))
    r_plain = b.align(guesses)
    b.set_params(R.make_params(anderson_memory=5, **work))
    r1 = b.align(guesses)
    r2 = b.align(guesses)
    it_plain = sum(r["iterations"] for r in r_plain)
    it_acc = sum(r["iterations"] for r in r1)
    d_plain = np.mean([synth.rmse_between(r_plain[p]["T"], plain[p]["T"], srcs[p]) for p in range(8)])
    d_acc = np.mean([synth.rmse_between(r1[p]["T"], plain[p]["T"], srcs[p]) for p in range(8)])
    print("iterations plain", it_plain, "accelerated", it_acc, "mean distance to the fixed point", d_plain, d_acc)
    assert it_acc < 0.8 * it_plain
    assert d_acc <= 1.5 * d_plain + 1e-4
    for x, y in zip(r1, r2):
        assert np.array_equal(x["T"], y["T"]) and x["iterations"] == y["iterations"]
    with pytest.raises(R.RtrError):
        b.set_params(R.make_params(anderson_memory=9, **work))


def test_cell_tuning_change_rebuilds_reused_grids():
    """With grids reused across aligns (rebuild-always off), a new cell
    factor takes effect at the next align: the grid's cell edge changes and
    the results equal an engine that rebuilds on every align; a later change
    of iteration settings only keeps that cell edge."""
    srcs, tgts, gts = synth.config3_batch(3, seed_offset=980, ns_range=(2000, 4000), nt_range=(5000, 9000))
    p = R.make_params(max_iterations=15, max_correspondence_distance=0.05, transformation_epsilon=1e-8)
    a, b = R.BatchRegistration(), R.BatchRegistration()
    for e in (a, b):
        e.set_clouds(srcs, tgts)
        e.set_params(p)
        e.set_problems(range(3), range(3))
    b.set_rebuild(False)
    b.align(gts)
    cell0 = b.grid(0)[0]
    for e in (a, b):
        e.set_tuning(cell_factor=0.6)
    rb, ra = b.align(gts), a.align(gts)
    assert b.grid(0)[0] != cell0
    for x, y in zip(ra, rb):
        assert np.array_equal(x["T"], y["T"]) and x["iterations"] == y["iterations"]
    cell1 = b.grid(0)[0]
    b.set_params(R.make_params(max_iterations=20, max_correspondence_distance=0.05, transformation_epsilon=1e-8))
    b.align(gts)
    assert b.grid(0)[0] == cell1


@pytest.mark.parametrize("mode", [0, 1])
def test_seeded_search_last_correspondences_bitexact(mode):
    """After the first iteration each query's search starts from its last
    nearest neighbour (a seed that trims the walk).  The correspondences of the
    last iteration a launch ran, as the seeded search found them, equal brute
    force at the transform that iteration queried with, bit for bit: near and
    far (multi-start) poses, launches stopped after 2, 4 and 7 iterations, both
    search variants, point-to-point and point-to-plane."""
    srcs, tgts, gts = synth.config3_batch(3, seed_offset=610, ns_range=(2000, 6000), nt_range=(8000, 20000),
                                          max_deg=10.0, max_t=0.05)
    starts = [multistart.so3_start_poses(s.astype(np.float64).mean(0), t.astype(np.float64).mean(0), 4)
              for s, t in zip(srcs, tgts)]
    guesses = [T for k in range(3) for T in [gts[k]] + list(starts[k][:2])]
    ids = np.repeat(np.arange(3), 3)
    for bpc in ((0, 1) if mode == 0 else (0,)):     # 4-deep compacted walk, 8-deep walk
        b = R.BatchRegistration()
        b.set_clouds(srcs, tgts)
        if mode == 1:
            b.compute_target_normals(20)
        b.set_problems(ids, ids)
        b.set_tuning(item_pts=256, blocks_per_cu=bpc, use_tile=0)
        for iters in (2, 4, 7):
            b.set_params(R.make_params(max_iterations=iters, max_correspondence_distance=0.05,
                                       transformation_epsilon=0.0, mode=mode))
            b.set_profiling(True)
            res = b.align(guesses)
            searched = b.timing()["queries_searched"]
            assert 0 < searched <= sum(len(srcs[i]) for i in ids) * iters
            for p in range(len(ids)):
                it = res[p]["iterations"]
                assert it >= 2
                rows = b.history(p, iters)[it - 1]
                T = np.eye(4)
                T[:3] = rows
                idx, d2 = b.last_correspondences(p)
                q = oracle.transform(T, srcs[ids[p]])
                i_ref, d_ref = oracle.nn_brute(q, tgts[ids[p]], 0.05)
                assert np.array_equal(idx, i_ref), f"{iters} iterations, problem {p}: {np.sum(idx != i_ref)} mismatches"
                assert np.array_equal(d2, d_ref)
            print(mode, bpc, iters, searched)


def test_surplus_workgroups_leaving_early_keeps_results():
    """Workgroups beyond one per CU leaving once few problems are unfinished
    (rtr_batch_set_retire) change who processes the remaining work items, not
    what they compute: the results are bit-identical."""
    srcs, tgts, gts = synth.config3_batch(8, seed_offset=700, ns_range=(2000, 8000), nt_range=(8000, 30000),
                                          max_deg=20.0, max_t=0.05)
    b = R.BatchRegistration()
    b.set_clouds(srcs, tgts)
    b.set_problems(np.arange(8), np.arange(8))
    b.set_params(R.make_params(max_iterations=30, max_correspondence_distance=0.05, transformation_epsilon=1e-8))
    b.set_tuning(item_pts=512, blocks_per_cu=0, use_tile=0, retire=0.0)
    ref = b.align()
    for frac in (0.25, 0.5, 1.0):
        b.set_tuning(retire=frac)
        assert b.tuning()[5] == frac
        res = b.align()
        for p in range(8):
            assert np.array_equal(res[p]["T"], ref[p]["T"]), (frac, p)
            assert res[p]["iterations"] == ref[p]["iterations"]


def test_coarse_to_fine_ends_at_a_full_resolution_fixed_point():
    """Coarse-to-fine iterations (rtr_params.coarse_stride): the first ones use
    every 4th source point (at most half of the iteration limit), only
    full-resolution ones can converge.  The last
    iteration's correspondences cover every source point, bit-exact against
    brute force, and one more plain FP64 iteration from the result moves it
    by no more than the transformation criterion allows (point-to-point,
    plain and accelerated; point-to-plane on the 8-deep search path: its
    correspondences checked the same way)."""
    srcs, tgts, gts = synth.config3_batch(6, seed_offset=800, ns_range=(3000, 8000), nt_range=(10000, 30000),
                                          max_deg=10.0, max_t=0.05)
    b = R.BatchRegistration()
    b.set_clouds(srcs, tgts)
    b.compute_target_normals(20)
    b.set_problems(np.arange(6), np.arange(6))
    for mode, aa in ((0, 0), (0, 5), (1, 0)):
        b.set_params(R.make_params(max_iterations=200, max_correspondence_distance=0.05, transformation_epsilon=1e-8,
                                   anderson_memory=aa, coarse_stride=4, coarse_iterations=50, mode=mode))
        b.set_tuning(use_tile=0)
        res = b.align()
        for p in range(6):
            r = res[p]
            assert r["state"] in (R._lib.CONV_TRANSFORM, R._lib.CONV_ABS_MSE), (aa, p, r["state"], r["iterations"])
            it = r["iterations"]
            rows = b.history(p, 200)[it - 1]
            T = np.eye(4)
            T[:3] = rows
            idx, d2 = b.last_correspondences(p)
            i_ref, d_ref = oracle.nn_brute(oracle.transform(T, srcs[p]), tgts[p], 0.05)
            assert np.array_equal(idx, i_ref) and np.array_equal(d2, d_ref), (mode, aa, p)
            assert 0 < r["coarse_iterations"] < it
            if mode == 0:
                nxt = oracle.icp(srcs[p], tgts[p], oracle.params(max_iter=1, r_max=0.05), guess=r["T"])
                print(aa, p, it, r["coarse_iterations"], synth.rmse_between(nxt["T"], r["T"], srcs[p]))
                assert synth.rmse_between(nxt["T"], r["T"], srcs[p]) < 2e-4
