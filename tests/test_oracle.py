"""The CPU oracle pinned against closed forms, brute force, library routines
and the reference's own data files (index-aligned input/output pairs of
earlier reference runs).  CPU only."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from helpers import affine_fit, golden, rot_err, trans_err
from pcd_np import read_xyz
from realtime_robot_amd import synth


def rigid(rng, deg=40.0, t=0.5):
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec(rng.normal(size=3) / np.sqrt(3) * np.deg2rad(deg)).as_matrix()
    T[:3, 3] = rng.uniform(-t, t, 3)
    return T


# ---------------------------------------------------------------- nearest neighbour

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_kdtree_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    t = rng.random((3000, 3)).astype(np.float32)
    q = (rng.random((800, 3)) * 1.4 - 0.2).astype(np.float32)
    kd = oracle.KdTree(t)
    for r in (None, 0.02, 0.1):
        i1, d1 = oracle.nn_brute(q, t, r)
        i2, d2 = kd.nn(q, r)
        assert np.array_equal(i1, i2) and np.array_equal(d1, d2)


@pytest.mark.parametrize("seed", [0, 1])
def test_kdtree_bounded_search_equals_unbounded(seed):
    """The radius-bounded search (used by the FP64 ICP) gives the unbounded
    search's answers, in both tie modes, on ties-heavy grid data and with
    queries straddling the radius exactly."""
    rng = np.random.default_rng(seed)
    t = (np.round(rng.random((4000, 3)) * 25) / 25).astype(np.float32)
    t = np.concatenate([t, t[:500]])                       # duplicate points: exact ties
    q = (np.round(rng.random((3000, 3)) * 50) / 50 * 1.3 - 0.15).astype(np.float32)
    kd = oracle.KdTree(t)
    for r in (0.02, 0.04, 0.1, 1.0):
        for tie in (True, False):
            i1, d1 = kd.nn(q, r, tie_lowest=tie)
            i2, d2 = kd.nn(q, r, tie_lowest=tie, bounded=True)
            assert np.array_equal(i1, i2) and np.array_equal(d1, d2), (r, tie)
            if tie:
                ib, db = oracle.nn_brute(q, t, r)
                assert np.array_equal(ib, i2) and np.array_equal(db, d2), r
    # a query at exactly the radius from its only neighbour is kept
    r = 0.5
    t1 = np.array([[0, 0, 0]], np.float32)
    x = np.sqrt(np.float32(oracle.r2_to_float(r * r)))
    qs = np.array([[x, 0, 0]], np.float32)
    d_exact = np.float32(qs[0, 0] * qs[0, 0])
    kd1 = oracle.KdTree(t1)
    for tie in (True, False):
        a = kd1.nn(qs, r, tie_lowest=tie)
        b = kd1.nn(qs, r, tie_lowest=tie, bounded=True)
        assert np.array_equal(a[0], b[0]) and (a[0][0] == 0) == (float(d_exact) <= r * r)


def test_brute_force_is_the_textbook_argmin():
    rng = np.random.default_rng(3)
    t = rng.random((500, 3)).astype(np.float32)
    q = rng.random((200, 3)).astype(np.float32)
    idx, d2 = oracle.nn_brute(q, t)
    # float32 ((dx*dx + dy*dy) + dz*dz), first minimum in index order
    d = q[:, None, :] - t[None, :, :]
    s = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    assert s.dtype == np.float32
    assert np.array_equal(idx, np.argmin(s, axis=1))
    assert np.array_equal(d2, s[np.arange(len(q)), idx])


def test_ties_go_to_lowest_index_and_nonfinite_targets_are_skipped():
    t = np.array([[1, 0, 0], [0, 1, 0], [1, 0, 0], [np.nan, 0, 0], [0, 0, 0.5]], np.float32)
    q = np.array([[0, 0, 0], [0.5, 0.5, 0], [np.nan, 0, 0]], np.float32)
    idx, d2 = oracle.nn_brute(q, t)
    assert idx[0] == 4 and d2[0] == np.float32(0.25)
    assert idx[1] == 0  # equidistant to 0, 1 and 2
    assert idx[2] == -1 and np.isinf(d2[2])
    kd = oracle.KdTree(t)
    i2, _ = kd.nn(q)
    assert np.array_equal(idx, i2)


def test_radius_rejection_is_d2_greater_than_rmax_squared():
    t = np.array([[0, 0, 0]], np.float32)
    r = 0.05
    r2f = oracle.r2_to_float(r * r)
    assert float(r2f) <= r * r and float(np.nextafter(r2f, np.float32(1))) > r * r
    # queries straddling the radius: accepted exactly when (double)d2 <= r*r
    x = np.float32(0.05) + (np.arange(-40, 41) * np.float32(1e-9)).astype(np.float32)
    q = np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1).astype(np.float32)
    idx, d2 = oracle.nn_brute(q, t, r)
    d2_exp = x * x
    accept = d2_exp.astype(np.float64) <= r * r
    assert accept.any() and (~accept).any()
    assert np.array_equal(idx == 0, accept)


def test_knn_is_lexicographic():
    rng = np.random.default_rng(4)
    t = rng.random((600, 3)).astype(np.float32)
    kd = oracle.KdTree(t)
    idx, d2 = kd.knn(t[:50], 20)
    d = t[:50, None, :] - t[None, :, :]
    s = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    order = np.lexsort((np.broadcast_to(np.arange(600), s.shape), s), axis=1)[:, :20]
    assert np.array_equal(idx, order)


# ---------------------------------------------------------------- small solves

def test_svd3_reconstructs_and_is_ordered():
    rng = np.random.default_rng(5)
    for _ in range(50):
        A = rng.normal(size=(3, 3))
        U, S, V = oracle.svd3(A)
        assert np.allclose(U @ np.diag(S) @ V.T, A, atol=1e-12)
        assert np.allclose(U.T @ U, np.eye(3), atol=1e-12) and np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert S[0] >= S[1] >= S[2] >= 0
        assert np.allclose(S, np.linalg.svd(A, compute_uv=False), atol=1e-12)


@pytest.mark.parametrize("seed", range(5))
def test_umeyama_recovers_exact_rigid_motion(seed):
    rng = np.random.default_rng(seed)
    T = rigid(rng, 170.0, 3.0)
    a = rng.normal(size=(200, 3)) + 5.0
    b = a @ T[:3, :3].T + T[:3, 3]
    assert np.allclose(oracle.umeyama(a, b), T, atol=1e-12)


def test_umeyama_reflection_guard_on_planar_points():
    rng = np.random.default_rng(9)
    a = np.concatenate([rng.random((100, 2)), np.zeros((100, 1))], 1)
    T = rigid(rng, 120.0)
    b = a @ T[:3, :3].T + T[:3, 3]
    R = oracle.umeyama(a, b)
    assert np.isclose(np.linalg.det(R[:3, :3]), 1.0)
    assert np.allclose(R, T, atol=1e-10)


def test_eig3_against_numpy():
    rng = np.random.default_rng(6)
    for _ in range(30):
        M = rng.normal(size=(3, 3))
        A = M @ M.T
        w, V = oracle.eig3(A)
        assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-12)
        assert np.allclose(A @ V, V * w, atol=1e-10)


def test_cholesky6_against_numpy_solve():
    rng = np.random.default_rng(7)
    for _ in range(20):
        M = rng.normal(size=(6, 6))
        A = M @ M.T + 0.1 * np.eye(6)
        b = rng.normal(size=6)
        x, rc = oracle.cholesky6(A, b)
        assert rc == 0 and np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_cholesky6_zeroes_unconstrained_directions():
    A = np.diag([1.0, 2.0, 0.0, 3.0, 0.0, 5.0])
    b = np.array([1.0, 2.0, 7.0, 3.0, 9.0, 5.0])
    x, rc = oracle.cholesky6(A, b)
    assert rc == 1 and np.allclose(x, [1, 1, 0, 1, 0, 1])


def test_euler_small_angle_parameterisation():
    rng = np.random.default_rng(8)
    for _ in range(10):
        x = rng.normal(size=6) * 0.3
        an = rng.normal(size=3)
        T = oracle.euler_to_T(x, an)
        R = Rotation.from_euler("ZYX", [x[2], x[1], x[0]]).as_matrix()  # Rz(g) Ry(b) Rx(a)
        assert np.allclose(T[:3, :3], R, atol=1e-14)
        # translation x[3:] acts in the anchor frame
        p = rng.normal(size=3)
        assert np.allclose(T[:3, :3] @ p + T[:3, 3], R @ (p - an) + x[3:] + an, atol=1e-12)


# ---------------------------------------------------------------- reference fixtures

def test_transform_fixture_rigid_rz90():
    """mcloudYasuo.pcd -> transformed_cloud2.pcd: Rz(90 deg) + t(1.5, 2.26, 0.05)."""
    a = read_xyz(golden("mcloudYasuo.pcd"))
    b = read_xyz(golden("transformed_cloud2.pcd"))
    T, res = affine_fit(a, b)
    assert res < 1e-7
    assert np.allclose(T[:3, :3], Rotation.from_euler("z", 90, degrees=True).as_matrix(), atol=1e-6)
    assert np.allclose(T[:3, 3], [1.5, 2.26, 0.05], atol=1e-6)
    out = oracle.transform(T, a)
    assert np.max(np.abs(out - b)) < 1.5e-7 * max(1.0, float(np.abs(b).max()))
    # the same pair as a Kabsch known answer with given correspondences
    K = oracle.umeyama(a.astype(np.float64), b.astype(np.float64))
    assert rot_err(K, T) < 1e-6 and trans_err(K, T) < 1e-6


def test_transform_fixture_nonrigid_coarse_stage_output():
    """T0_m8111.pcd -> transformed_cloud3.pcd: the coarse stage's non-rigid
    3x3 (scale x 0.8999 on y, z and Rz(40 deg)) applied by transformPointCloud."""
    a = read_xyz(golden("T0_m8111.pcd"))
    b = read_xyz(golden("transformed_cloud3.pcd"))
    T, res = affine_fit(a, b)
    assert res < 1e-7
    # rows 1-2 carry the 0.8999 scale; row 0 of diag(1, s, s) Rz(40) is (cos, -sin, 0)
    ang = np.degrees(np.arctan2(-T[0, 1], T[0, 0]))
    assert abs(ang - 40.0) < 1e-3
    assert abs(np.linalg.norm(T[1, :3]) - 0.8999) < 1e-3 and abs(np.linalg.norm(T[0, :3]) - 1.0) < 1e-4
    out = oracle.transform(T, a)
    assert np.max(np.abs(out - b)) < 2e-7 * max(1.0, float(np.abs(b).max()))


def test_transform_fixture_uninitialised_matrix_ignores_row_four():
    """5.pcd -> transformed_cloud.pcd: every entry of the top three rows is the
    MSVC debug fill 0xCCCCCCCC (-107374176.0f); the bottom row is unused."""
    a = read_xyz(golden("5.pcd"))
    b = read_xyz(golden("transformed_cloud.pcd"))
    cc = np.frombuffer(np.uint32(0xCCCCCCCC).tobytes(), np.float32)[0]
    T = np.full((4, 4), float(cc))
    T[3] = [1e30, -7.0, 3.0, 42.0]  # anything: row 4 is not read
    out = oracle.transform(T, a)
    rel = np.abs(out.astype(np.float64) - b) / np.maximum(np.abs(b.astype(np.float64)), 1.0)
    assert rel.max() < 1e-7


def test_kabsch_known_answer_pure_translation():
    """70761.pcd -> 70761_c.pcd is x + 1, y + 0.5 (function.h:137-143)."""
    a = read_xyz(golden("70761.pcd"))
    b = read_xyz(golden("70761_c.pcd"))
    K = oracle.umeyama(a.astype(np.float64), b.astype(np.float64))
    assert rot_err(K, np.eye(4)) < 1e-6 and np.allclose(K[:3, 3], [1.0, 0.5, 0.0], atol=1e-6)


def test_icp_known_answer_pure_translation():
    """5.pcd -> mcloud.pcd is a translation (0.8, 0.8, 0): plain ICP from the
    identity with PCL's unbounded correspondence distance recovers it."""
    a = read_xyz(golden("5.pcd"))
    b = read_xyz(golden("mcloud.pcd"))
    r = oracle.icp(a, b, oracle.params(max_iter=50))
    assert np.allclose(r["T"][:3, 3], [0.8, 0.8, 0.0], atol=1e-6)
    assert rot_err(r["T"], np.eye(4)) < 1e-6
    assert r["mse"] < 1e-12 and r["converged"]


def test_icp_from_identity_fails_on_the_shifted_chair():
    """70761 -> 70761_c: plain ICP from the identity falls into a wrong
    minimum, the case the learned proposer and multi-start exist for."""
    a = read_xyz(golden("70761.pcd"))
    b = read_xyz(golden("70761_c.pcd"))
    truth = np.eye(4)
    truth[:3, 3] = [1.0, 0.5, 0.0]
    r = oracle.icp(a, b, oracle.params(max_iter=50))
    assert synth.rmse_between(r["T"], truth, a) > 0.05


# ---------------------------------------------------------------- ICP loop semantics

def test_icp_mse_is_non_increasing_with_unbounded_distance():
    src, tgt, _ = synth.config1_pair(4)
    r = oracle.icp(src, tgt, oracle.params(max_iter=15), trace=True)
    f = [oracle.fitness(src, tgt, np.eye(4))] + [oracle.fitness(src, tgt, T) for T in r["trace"]]
    assert all(f[i + 1] <= f[i] * (1 + 1e-9) for i in range(len(f) - 1))


def test_icp_identity_input_converges_immediately():
    """Source == target: the first increment is the identity up to rounding;
    with epsilon 0 the transform test needs cos = 1 exactly, so the absolute
    MSE test (|0 - 0| < 1e-12) stops the second iteration at the latest."""
    rng = np.random.default_rng(10)
    a = rng.random((500, 3)).astype(np.float32)
    r = oracle.icp(a, a, oracle.params(max_iter=10))
    assert r["state"] in (oracle.CONV_TRANSFORM, oracle.CONV_ABS_MSE) and r["iterations"] <= 2
    assert np.allclose(r["T"], np.eye(4), atol=1e-12) and r["mse"] == 0.0
    r = oracle.icp(a, a, oracle.params(max_iter=10, eps_T=1e-12))
    assert r["state"] == oracle.CONV_TRANSFORM and r["iterations"] == 1


def test_icp_max_iterations_and_too_few_correspondences():
    src, tgt, _ = synth.config1_pair(5)
    r = oracle.icp(src, tgt, oracle.params(max_iter=3, r_max=0.05))
    assert r["iterations"] == 3 and r["state"] == oracle.CONV_ITERATIONS and r["converged"]
    r = oracle.icp(src + np.float32(5), tgt, oracle.params(max_iter=3, r_max=0.05))
    assert r["iterations"] == 0 and r["state"] == oracle.FAIL_NO_CORR and not r["converged"]


def test_icp_relative_mse_criterion():
    src, tgt, _ = synth.config1_pair(6)
    r = oracle.icp(src, tgt, oracle.params(max_iter=100, r_max=0.05, eps_fit=1e-2))
    assert r["state"] == oracle.CONV_REL_MSE and r["iterations"] < 100


def test_point_to_plane_recovers_small_motion():
    rng = np.random.default_rng(11)
    # three orthogonal planes: the 6x6 system is well conditioned
    n = 3000
    pts = np.zeros((n, 3))
    nrm = np.zeros((n, 3))
    for k in range(3):
        sl = slice(k * n // 3, (k + 1) * n // 3)
        m = sl.stop - sl.start
        p = rng.random((m, 3))
        p[:, k] = 0.0
        pts[sl] = p
        nrm[sl, k] = 1.0
    T = rigid(rng, 2.0, 0.01)
    src = ((pts - T[:3, 3]) @ T[:3, :3]).astype(np.float32)  # T^-1 * pts
    r = oracle.icp(src, pts.astype(np.float32), oracle.params(max_iter=30, mode=1, anchor=(0.5, 0.5, 0.5)),
                   tgt_normals=nrm.astype(np.float32))
    assert rot_err(r["T"], T) < 1e-5 and trans_err(r["T"], T) < 1e-5


def test_fitness_score_squared_vs_unsquared_range_quirk():
    t = np.array([[0, 0, 0]], np.float32)
    s = np.array([[0.1, 0, 0], [0.3, 0, 0]], np.float32)
    # d2 = 0.01 and 0.09; max_range 0.05 (compared with d2, unsquared) keeps only the first
    assert np.isclose(oracle.fitness(s, t, np.eye(4), 0.05), 0.01)
    assert oracle.fitness(s, t, np.eye(4), 0.001) == np.finfo(np.float64).max


# ---------------------------------------------------------------- voxels and normals

def test_voxel_accumulate_matches_a_dictionary():
    pts = read_xyz(golden("ground.pcd"))
    rng = np.random.default_rng(12)
    T = rigid(rng, 30.0, 1.0)
    inv = np.float32(1.0 / np.float64(np.float32(0.02)))
    keys, sums, counts, dropped = oracle.voxel_accumulate(pts, inv, T)
    q = oracle.transform(T, pts)
    k3 = np.floor(q * inv).astype(np.int64)
    ref = {}
    for kk, p in zip(map(tuple, k3), q.astype(np.float64)):
        s, c = ref.get(kk, (np.zeros(3), 0))
        ref[kk] = (s + p, c + 1)
    assert dropped == 0 and len(keys) == len(ref) and counts.sum() == len(pts)
    bias = 1 << 20
    for kk, s, c in zip(keys, sums, counts):
        t3 = tuple(int(((kk >> (21 * a)) & 0x1FFFFF) - bias) for a in range(3))
        rs, rc = ref[t3]
        assert rc == c and np.allclose(rs, s, rtol=1e-12, atol=1e-9)
    # centroid lies inside its voxel
    cen = sums / counts[:, None]
    kc = np.stack([((keys >> (21 * a)) & 0x1FFFFF) - bias for a in range(3)], 1)
    assert np.all(cen >= kc * 0.02 - 1e-6) and np.all(cen <= (kc + 1) * 0.02 + 1e-6)


def test_normals_on_plane_and_sphere():
    rng = np.random.default_rng(13)
    plane = np.concatenate([rng.random((2000, 2)), np.full((2000, 1), 0.3)], 1).astype(np.float32)
    n, curv = oracle.normals(plane, 20, vp=(0.5, 0.5, 5.0))
    assert np.allclose(n, [0, 0, 1], atol=1e-5)
    assert np.all(curv < 1e-6)
    v = rng.normal(size=(4000, 3))
    sph = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    n, _ = oracle.normals(sph, 20, vp=(0, 0, 0))
    cosang = np.sum(n * -sph, axis=1)  # flipped towards the centre viewpoint
    assert np.median(cosang) > 0.999 and np.min(cosang) > 0.98


def test_voxel_downsample_matches_grouping():
    pts = read_xyz(golden("chair1.pcd"))
    out, counts = oracle.voxel_downsample(pts, 0.01)
    inv = np.float32(1.0 / np.float64(np.float32(0.01)))
    k = np.floor(pts * inv).astype(np.int64)
    lo = k.min(0)
    d = k.max(0) - lo + 1
    lin = (k[:, 0] - lo[0]) + (k[:, 1] - lo[1]) * d[0] + (k[:, 2] - lo[2]) * d[0] * d[1]
    u, inverse = np.unique(lin, return_inverse=True)
    ref = np.zeros((len(u), 3))
    np.add.at(ref, inverse, pts.astype(np.float64))
    ref /= np.bincount(inverse)[:, None]
    assert len(out) == len(u) and np.allclose(out, ref, atol=1e-7)


# ---------------------------------------------------------------- PCL-semantics baseline

def test_pcl_semantics_baseline_agrees_with_fp64_oracle():
    src, tgt, _ = synth.config1_pair(7)
    p = oracle.params(max_iter=50, r_max=0.05, eps_T=1e-8)
    a = oracle.pcl_icp(src, tgt, p)
    b = oracle.icp(src, tgt, p)
    assert synth.rmse_between(a["T"], b["T"], src) < 1e-3
    batch = oracle.pcl_icp_batch([src, src], [tgt, tgt], p, nthreads=2)
    assert np.array_equal(batch[0]["T"], a["T"]) and np.array_equal(batch[1]["T"], a["T"])


def test_tdf_oracle_matches_scipy_edt():
    """The brute-force TDF (kernel.cu:9-31 semantics, bounds fixed) equals
    scipy's exact Euclidean distance transform, squared and capped."""
    from scipy import ndimage
    rng = np.random.default_rng(31)
    dims = (13, 9, 11)
    occ = rng.integers(0, [13, 9, 11], size=(25, 3))
    out = oracle.tdf(occ, dims, 30.0)
    grid = np.ones((dims[2], dims[1], dims[0]), bool)
    grid[occ[:, 2], occ[:, 1], occ[:, 0]] = False
    ref = np.minimum(ndimage.distance_transform_edt(grid) ** 2, 30.0)
    assert np.array_equal(out, ref.astype(np.float32))
    # out-of-grid voxels are ignored; an empty grid is the cap everywhere
    assert np.array_equal(oracle.tdf([[-1, 0, 0], [13, 0, 0]], dims, 900.0), np.full(out.shape, 900.0, np.float32))
    one = oracle.tdf([[2, 3, 4]], (5, 5, 5), 900.0)
    assert one[4, 3, 2] == 0 and one[0, 0, 0] == 4 + 9 + 16
