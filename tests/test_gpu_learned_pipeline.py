"""The learned-init pipeline behind the C ABI (rtr_learned_*, pipeline.cpp).

1. Bit parity with the same pipeline composed from the library's primitives
   in Python (the network, a refine engine, the judge, a fallback engine and
   the best-start rule), through ctypes: proposals, refined poses,
   iterations, judge outputs and rejections identical.
2. Failure injection (the fallback that the reference's consensus plays when
   no hypothesis agrees, function.h:35-109): proposals of chosen pairs are
   replaced by poses 90-180 degrees wrong.  The judge must reject exactly
   those pairs (plus any it rejected before), the multi-start fallback must
   bring them within 1 mm of the FP64 oracle's registration from the
   generating pose, and every other pair must come out bit-identical.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from realtime_robot_amd import learned, multistart, synth
from realtime_robot_amd.registration import BatchRegistration

pytestmark = pytest.mark.gpu

N = 24


@pytest.fixture(scope="module")
def pairs():
    srcs, tgts, gts = synth.config4_batch(N)
    cs = np.stack([s.astype(np.float64).mean(0) for s in srcs])
    ct = np.stack([t.astype(np.float64).mean(0) for t in tgts])
    return srcs, tgts, gts, cs, ct


@pytest.fixture(scope="module")
def pipeline(pairs):
    srcs, tgts, _, cs, ct = pairs
    lr = learned.LearnedRegistration(max_pairs=N)
    lr.set_clouds(srcs, tgts, cs, ct)
    return lr


def _compose(pairs, weights, proposals=None):
    """The pipeline from the primitives: Proposer + refine BatchRegistration
    + judge + fallback BatchRegistration + multistart.select_best."""
    srcs, tgts, _, cs, ct = pairs
    params = learned.default_params()
    rots = multistart.so3_rotations(8)
    if proposals is None:
        prop = learned.Proposer(max_pairs=N * 8, weights=weights)
        prop.set_clouds(srcs, tgts)
        _, _, proposals = prop.propose_multi(rots, 0.1)
    refine = learned.L.Params.from_buffer_copy(params)
    refine.anderson_memory, refine.coarse_stride, refine.coarse_iterations = 5, 4, 25
    reg = BatchRegistration()
    reg.set_params(refine)
    reg.set_clouds(srcs, tgts)
    reg.set_problems(range(N), range(N))
    reg.set_rebuild(False)
    reg.build()
    res = reg.align(proposals)
    rmse, frac, _ = reg.judge(20, 1024, 0.02)
    fail = np.where((rmse > 2e-3) | (frac < 0.8))[0]
    inits = np.array(proposals, np.float64, copy=True)
    if len(fail):
        fb = BatchRegistration()
        fb.set_params(params)
        fb.set_tuning(item_pts=256, blocks_per_cu=1)
        fb.set_graph(False)
        cf, cpp = reg.tuning()[:2]
        fb.set_tuning(cell_factor=cf, cells_per_point=cpp)
        fb.set_clouds([srcs[i] for i in fail], [tgts[i] for i in fail])
        ids = np.repeat(np.arange(len(fail)), 8)
        fb.set_problems(ids, ids)
        g = np.stack([T for i in fail for T in multistart.start_poses_from(rots, cs[i], ct[i])])
        r8 = fb.align(g)
        for k, i in enumerate(fail):
            b = multistart.select_best(r8[k * 8:(k + 1) * 8])
            res[i] = r8[k * 8 + b]
            inits[i] = g[k * 8 + b]
    return np.asarray(proposals), res, rmse, frac, fail, inits


def _same(a, b, label):
    for i in range(N):
        assert np.array_equal(a[i]["T"], b[i]["T"]), f"{label}: pair {i} transforms differ"
        assert a[i]["iterations"] == b[i]["iterations"] and a[i]["state"] == b[i]["state"], f"{label}: pair {i}"


def test_pipeline_matches_composition_bit_for_bit(pairs, pipeline):
    weights, _ = learned.load_weights()
    res = pipeline.align()
    st = pipeline.stats
    props, ref, rmse, frac, fail, inits = _compose(pairs, weights)
    assert np.array_equal(st["proposals"], props)
    _same(res, ref, "align")
    assert np.array_equal(st["judge_rmse"], rmse) and np.array_equal(st["judge_frac"], frac)
    assert np.array_equal(st["fallback_pairs"], fail)
    assert np.array_equal(st["inits"], inits)
    print(f"pipeline: {N} pairs, {len(fail)} rejected, {st['ms_total']:.1f} ms", flush=True)


def _fp64_from(pairs, which, poses):
    srcs, tgts = pairs[0], pairs[1]
    p = oracle.params(max_iter=50, r_max=0.05, eps_T=1e-8)
    with ThreadPoolExecutor(8) as ex:
        return dict(zip(which, ex.map(lambda i: oracle.icp(srcs[i], tgts[i], p, guess=poses[i])["T"], which)))


def test_fallback_rescues_injected_wrong_proposals(pairs, pipeline):
    srcs, tgts, gts, cs, ct = pairs
    base = pipeline.align()
    props = pipeline.stats["proposals"].copy()
    rejected0 = set(pipeline.stats["fallback_pairs"].tolist())
    # pairs the plain multi-start itself registers (the fallback can only
    # recover what its starts reach): the 8-start result within 1 mm of the
    # FP64 registration from the generating pose
    ms = learned.MultiStartRegistration(learned.default_params(), n_starts=8)
    ms.set_clouds(srcs, tgts, cs, ct)
    r_ms = ms.align()
    cand = [i for i in range(N) if i not in rejected0]
    ref = _fp64_from(pairs, cand, gts)
    good = [i for i in cand if synth.rmse_between(r_ms[i]["T"], ref[i], srcs[i]) <= 1e-3]
    assert len(good) >= 2, f"multi-start registers only {good}"
    inject = good[:4]
    rng = np.random.default_rng(42)
    bad = props.copy()
    for i in inject:
        # turn the proposal by 90-180 degrees about a random axis through the
        # source centroid's image
        D = np.eye(4)
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        th = np.deg2rad(rng.uniform(90.0, 180.0))
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        D[:3, :3] = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * (K @ K)
        c = props[i][:3, :3] @ cs[i] + props[i][:3, 3]
        D[:3, 3] = c - D[:3, :3] @ c
        bad[i] = D @ props[i]
    res = pipeline.align_from(bad)
    st = pipeline.stats
    assert set(st["fallback_pairs"].tolist()) == rejected0 | set(inject), (st["fallback_pairs"], inject)
    for i in inject:
        e = synth.rmse_between(res[i]["T"], ref[i], srcs[i])
        assert e <= 1e-3, f"pair {i}: fallback result {e * 1e3:.2f} mm from the FP64 registration"
    for i in range(N):
        if i not in rejected0 and i not in inject:
            assert np.array_equal(res[i]["T"], base[i]["T"]), f"pair {i} changed"
    # and the composition from the primitives agrees bit for bit, fallback included
    _, ref_res, _, _, fail, inits = _compose(pairs, None, proposals=bad)
    _same(res, ref_res, "align_from")
    assert np.array_equal(st["fallback_pairs"], fail) and np.array_equal(st["inits"], inits)
    print(f"injected {inject}: rejected {st['fallback_pairs'].tolist()}, fallback {st['ms_fallback']:.1f} ms",
          flush=True)


def test_cpp_learned_registration_drop_in(tmp_path, pairs):
    """The C++ class (rtr::LearnedRegistration, library defaults: its own
    rotation sets, FP64 centroids) on PCD files, weights from the raw float32
    file: the registrations it accepts land within 1 mm of the FP64 oracle's
    registration from the generating pose for most pairs, as the Python
    pipeline's do."""
    import os
    import subprocess

    import realtime_robot_amd as R
    from realtime_robot_amd import _lib as L
    srcs, tgts, gts = pairs[0][:8], pairs[1][:8], pairs[2][:8]
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "learned_batch"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "learned_batch.cpp"), "-o", str(exe),
                    "-L", os.path.dirname(L.LIB_PATH), "-lrtr", "-Wl,-rpath," + os.path.dirname(L.LIB_PATH)],
                   check=True)
    args = [str(exe), learned.WEIGHTS_F32]
    clouds = []
    for i in range(len(srcs)):
        for name, c in (("s", srcs[i]), ("t", tgts[i])):
            path = tmp_path / f"{name}{i}.pcd"
            R.save_pcd_ascii(str(path), c)
            args.append(str(path))
        clouds.append((R.load_pcd(str(tmp_path / f"s{i}.pcd"))[0], R.load_pcd(str(tmp_path / f"t{i}.pcd"))[0]))
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("pair ")]
    assert len(lines) == len(srcs)
    ref = _fp64_from((list(c[0] for c in clouds), list(c[1] for c in clouds)), range(len(srcs)), gts)
    ok = 0
    for i, ln in enumerate(lines):
        T = np.array([float(v) for v in ln[7:23]]).reshape(4, 4).T
        ok += synth.rmse_between(T, ref[i], clouds[i][0]) <= 1e-3
    print(r.stdout.splitlines()[-1], f"; {ok}/{len(srcs)} within 1 mm", flush=True)
    assert ok >= len(srcs) - 2
