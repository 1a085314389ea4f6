"""Learned pose proposer and the learned-init registration pipeline.

The reference gets ICP's initial pose from its coarse stage: keypoints and
plane areas, a 10-degree TDF yaw search and an exhaustive consensus
(RealTimeRobot.cpp:45-105, matching.h:122-222, function.h:35-109).  Here a
point network proposes it instead:

  * ``Proposer``: the HIP network in librtr (rtr_prop_*): farthest-point
    sampling to 1024 points per cloud, shared MLP 3-64-128-256 + max-pool per
    cloud, head 512-256-9 (6D rotation + translation), BF16 MFMA with FP32
    accumulation, forward, backward and Adam on the device.
  * ``build_training_set`` / ``train``: canonical pairs sampled from random
    asymmetric box composites, labelled by the GPU ICP's converged pose (batch
    ICP started at the generating pose), farthest-point sampled on the device;
    every training pair is a base pair whose source is moved by a fresh random
    rigid motion (rotation U[0, 180 deg] about a uniform axis, translation
    U[-0.1, 0.1]^3 m), with the label composed accordingly.
  * ``LearnedRegistration``: propose -> batch ICP refine -> judge
    (rtr_batch_judge: RMS distance of the inliers to their k-NN PCA plane
    where the neighbourhood is planar, and the inlier fraction) -> 8-start ICP
    for the pairs whose RMSE is above 2 mm or whose inlier fraction is below
    0.8.
  * ``MultiStartRegistration``: the plain multi-start ICP it is compared with
    (and falls back to).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib as L
from . import multistart, synth
from .registration import BatchRegistration, make_params, result_dicts

KP = 1024
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "proposer.npz")
WEIGHTS_F32 = os.path.splitext(WEIGHTS_PATH)[0] + ".f32"

_SHAPES = [("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 128)), ("b2", (128,)), ("W3", (128, 256)), ("b3", (256,)),
           ("W4", (512, 256)), ("b4", (256,)), ("W5", (256, 9)), ("b5", (9,))]


def n_params() -> int:
    n = C.c_size_t()
    L.check(L.lib.rtr_prop_num_params(C.byref(n)))
    return n.value


def init_weights(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in; 3 real inputs for the first layer), zero
    biases; the last layer is scaled down so the first proposals stay near
    the centred identity."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in _SHAPES:
        if name.startswith("W"):
            fan_in = 3 if name == "W1" else shp[0]
            w = rng.normal(0.0, math.sqrt(2.0 / fan_in), size=shp)
            if name == "W5":
                w *= 0.1
            parts.append(w.reshape(-1))
        elif name == "b5":
            parts.append(np.array([1, 0, 0, 0, 1, 0, 0, 0, 0], np.float64))  # 6D identity, t = 0
        else:
            parts.append(np.zeros(shp).reshape(-1))
    w = np.concatenate(parts).astype(np.float32)
    assert len(w) == n_params()
    return w


def save_weights(path: str, w, meta: dict) -> None:
    """The weights as .npz (FP32 array + JSON training record) and, beside
    it, as a raw little-endian float32 file (same stem, .f32) for
    rtr_learned_load_weights."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, w=np.asarray(w, np.float32), meta=np.array(json.dumps(meta)))
    np.asarray(w, "<f4").tofile(os.path.splitext(path)[0] + ".f32")


def load_weights(path: str = WEIGHTS_PATH):
    """(weights float32, meta dict); plain arrays only (no pickle)."""
    with np.load(path, allow_pickle=False) as z:
        w = z["w"].astype(np.float32)
        meta = json.loads(str(z["meta"]))
    if len(w) != n_params():
        raise ValueError(f"{path}: {len(w)} parameters, the network has {n_params()}")
    return w, meta


class Proposer:
    """The device network (rtr_prop_*): FPS, forward, backward, Adam."""

    def __init__(self, max_pairs: int = 256, device: int = 0, weights=None):
        h = C.c_void_p()
        L.check(L.lib.rtr_prop_create(C.byref(h), int(device), int(max_pairs)))
        self._h = h
        self._owner = None
        self.max_pairs = max_pairs
        self.n_pairs = 0
        self.set_weights(init_weights(0) if weights is None else weights)

    @classmethod
    def borrow(cls, handle, owner, max_pairs: int):
        """A view of the network another handle owns (rtr_learned_batches)."""
        self = cls.__new__(cls)
        self._h = handle
        self._owner = owner
        self.max_pairs = max_pairs
        self.n_pairs = 0
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_owner", None) is None:
            L.lib.rtr_prop_destroy(h)
        self._h = None

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if len(w) != n_params():
            raise ValueError(f"expected {n_params()} parameters, got {len(w)}")
        L.check(L.lib.rtr_prop_set_weights(self._h, L.ptr(w)))

    def get_weights(self) -> np.ndarray:
        w = np.empty(n_params(), np.float32)
        L.check(L.lib.rtr_prop_get_weights(self._h, L.ptr(w)))
        return w

    def set_clouds(self, sources, targets):
        src = [L.pts3(s) for s in sources]
        tgt = [L.pts3(t) for t in targets]
        if len(src) != len(tgt):
            raise ValueError("sources and targets must pair up")
        sp = (C.c_void_p * len(src))(*[s.ctypes.data for s in src])
        tp = (C.c_void_p * len(tgt))(*[t.ctypes.data for t in tgt])
        ns = (C.c_size_t * len(src))(*[len(s) for s in src])
        nt = (C.c_size_t * len(tgt))(*[len(t) for t in tgt])
        L.check(L.lib.rtr_prop_set_clouds(self._h, len(src), sp, ns, tp, nt))
        self.n_pairs = len(src)

    def fps(self):
        """(points (2n, 1024, 3) centred, centroids (2n, 3), subsample indices (2n, 1024)); clouds src0, tgt0, ..."""
        n = 2 * self.n_pairs
        pts = np.empty((n, KP, 3), np.float32)
        cen = np.empty((n, 3), np.float32)
        sel = np.empty((n, KP), np.int32)
        L.check(L.lib.rtr_prop_fps(self._h, L.ptr(pts), L.ptr(cen), L.ptr(sel)))
        return pts, cen, sel

    def propose(self) -> np.ndarray:
        """(n, 4, 4) row-major poses mapping each source onto its target."""
        cm = np.empty((self.n_pairs, 16))
        L.check(L.lib.rtr_prop_propose(self._h, L.ptr(cm)))
        return np.ascontiguousarray(np.transpose(cm.reshape(-1, 4, 4), (0, 2, 1)))

    def propose_multi(self, rotations, tau: float = 0.1):
        """k hypotheses per pair from the source samples rotated by each of
        `rotations` (k, 3, 3): (poses (n, k, 4, 4), scores (n, k), best (n, 4, 4))."""
        R = np.ascontiguousarray(rotations, np.float64).reshape(-1, 9)
        k, n = len(R), self.n_pairs
        cm = np.empty((n, k, 16))
        sc = np.empty((n, k))
        best = np.empty((n, 16))
        L.check(L.lib.rtr_prop_propose_multi(self._h, k, L.ptr(R), float(tau), L.ptr(cm), L.ptr(sc), L.ptr(best)))
        poses = np.ascontiguousarray(np.transpose(cm.reshape(n, k, 4, 4), (0, 1, 3, 2)))
        return poses, sc, np.ascontiguousarray(np.transpose(best.reshape(n, 4, 4), (0, 2, 1)))

    def set_fused(self, on: bool):
        """Inference through the fused point-MLP + max-pool kernel (default)
        or the layer-by-layer products; the features are identical."""
        L.check(L.lib.rtr_prop_set_fused(self._h, int(on)))

    def last_ms(self) -> float:
        ms = C.c_float()
        L.check(L.lib.rtr_prop_last_ms(self._h, C.byref(ms)))
        return ms.value

    def forward_backward(self, xyz, labels=None, grad: bool = True) -> dict:
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0] // (2 * KP)
        O = np.empty((n, 9), np.float32)
        rt = np.empty((n, 12))
        out = {"O": O}
        lab = None if labels is None else np.ascontiguousarray(labels, np.float64).reshape(n, 12)
        loss = C.c_float()
        g = np.empty(n_params(), np.float32) if (lab is not None and grad) else None
        L.check(L.lib.rtr_prop_forward_backward(self._h, n, L.ptr(xyz), L.ptr(lab), L.ptr(O), L.ptr(rt),
                                                C.byref(loss) if lab is not None else None, L.ptr(g)))
        out["R"] = rt[:, :9].reshape(n, 3, 3)
        out["t"] = rt[:, 9:].copy()
        if lab is not None:
            out["loss"] = loss.value
        if g is not None:
            out["grad"] = g
        return out

    def train_step(self, xyz, labels, lr: float) -> float:
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        n = xyz.shape[0] // (2 * KP)
        lab = np.ascontiguousarray(labels, np.float64).reshape(n, 12)
        loss = C.c_float()
        L.check(L.lib.rtr_prop_train_step(self._h, n, L.ptr(xyz), L.ptr(lab), float(lr), C.byref(loss)))
        return loss.value


def centred_labels(T, cs, ct) -> np.ndarray:
    """(n, 4, 4) poses mapping source onto target -> (n, 12) labels in the
    centred frames: R row-major, then t* = R cs + t - ct."""
    T = np.asarray(T, np.float64)
    R = T[:, :3, :3]
    t = np.einsum("nij,nj->ni", R, np.asarray(cs, np.float64)) + T[:, :3, 3] - np.asarray(ct, np.float64)
    return np.concatenate([R.reshape(-1, 9), t], axis=1)


def random_rigid_batch(rng: np.random.Generator, n: int, max_deg: float = 180.0, max_t: float = 0.1):
    """(n, 4, 4): rotation U[0, max_deg] about a uniform axis, translation U[-max_t, max_t]^3."""
    R = multistart._rodrigues(rng.normal(size=(n, 3)), np.deg2rad(rng.uniform(0.0, max_deg, size=n)))
    T = np.tile(np.eye(4), (n, 1, 1))
    T[:, :3, :3] = R
    T[:, :3, 3] = rng.uniform(-max_t, max_t, size=(n, 3))
    return T


class TrainingSet:
    """Base pairs: centred FPS points of source and target, their centroids
    and the GPU-ICP label (canonical frame: source and target both in the
    piece frame, the label close to the identity)."""

    def __init__(self, src_pts, tgt_pts, cs, ct, T_label):
        self.src_pts, self.tgt_pts = src_pts, tgt_pts
        self.cs, self.ct, self.T_label = cs, ct, T_label

    def __len__(self):
        return len(self.cs)

    def batch(self, idx, rng: np.random.Generator, max_deg: float = 180.0, max_t: float = 0.1):
        """Network inputs (len(idx) * 2 * 1024, 3) and labels (len(idx), 12)
        for base pairs idx, each source moved by a fresh random rigid motion
        A: the source's centred FPS points rotate with A's rotation, its
        centroid maps to A cs, and the pose becomes T_label A^-1."""
        n = len(idx)
        A = random_rigid_batch(rng, n, max_deg, max_t)
        Q = A[:, :3, :3]
        xyz = np.empty((n, 2, KP, 3), np.float32)
        np.matmul(self.src_pts[idx], np.transpose(Q, (0, 2, 1)).astype(np.float32), out=xyz[:, 0])
        xyz[:, 1] = self.tgt_pts[idx]
        cs = np.einsum("nij,nj->ni", Q, self.cs[idx]) + A[:, :3, 3]
        T = self.T_label[idx] @ np.linalg.inv(A)
        return xyz.reshape(-1, 3), centred_labels(T, cs, self.ct[idx])


def build_training_set(n_base: int, seed: int = 20240, chunk: int = 512, device: int = 0, log=print,
                       ns_range=(2000, 6000), nt_range=(4000, 12000), noise: float = 0.001) -> TrainingSet:
    """n_base canonical pairs labelled by the GPU ICP (started at the
    generating pose: identity) and farthest-point sampled on the device."""
    prop = Proposer(max_pairs=chunk, device=device)
    reg = BatchRegistration(device)
    reg.set_params(default_params())
    S, Tp, CS, CT, TL = [], [], [], [], []
    kept = 0
    t0 = time.time()
    for srcs, tgts in synth.training_base(n_base, seed, ns_range=ns_range, nt_range=nt_range, noise=noise,
                                          chunk=chunk):
        m = len(srcs)
        reg.set_clouds(srcs, tgts)
        reg.set_problems(np.arange(m), np.arange(m))
        res = reg.align(None)
        T = np.stack([r["T"] for r in res])
        # a converged label stays within a few mm of the generating pose
        ok = np.array([r["state"] != L.FAIL_NO_CORRESPONDENCES and
                       synth.rmse_between(T[i], np.eye(4), srcs[i]) < 5e-3 for i, r in enumerate(res)])
        prop.set_clouds(srcs, tgts)
        pts, cen, _ = prop.fps()
        pts = pts.reshape(m, 2, KP, 3)
        cen = cen.reshape(m, 2, 3).astype(np.float64)
        S.append(pts[ok, 0])
        Tp.append(pts[ok, 1])
        CS.append(cen[ok, 0])
        CT.append(cen[ok, 1])
        TL.append(T[ok])
        kept += int(ok.sum())
        log(f"training set: {kept} pairs labelled ({time.time() - t0:.1f} s)")
    return TrainingSet(np.concatenate(S), np.concatenate(Tp), np.concatenate(CS), np.concatenate(CT),
                       np.concatenate(TL))


def rotation_error_deg(Ra, Rb) -> np.ndarray:
    c = (np.einsum("nij,nij->n", Ra, Rb) - 1.0) / 2.0
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


def train(ts: TrainingSet, steps: int, batch: int = 256, lr: float = 1e-3, lr_min: float = 1e-5,
          warmup: int = 200, seed: int = 0, weights=None, device: int = 0, log=print, log_every: int = 500):
    """Adam on the device with a warm-up + cosine learning rate.  Returns
    (weights, loss history)."""
    from concurrent.futures import ThreadPoolExecutor
    prop = Proposer(max_pairs=batch, device=device, weights=init_weights(seed) if weights is None else weights)
    rng = np.random.default_rng(seed + 1)
    state = {"order": rng.permutation(len(ts)), "pos": 0}

    def next_batch():
        if state["pos"] + batch > len(state["order"]):
            state["order"] = rng.permutation(len(ts))
            state["pos"] = 0
        idx = state["order"][state["pos"]:state["pos"] + batch]
        state["pos"] += batch
        return ts.batch(idx, rng)

    hist = []
    t0 = time.time()
    pool = ThreadPoolExecutor(1)          # builds the next batch while the device trains
    fut = pool.submit(next_batch)
    for step in range(steps):
        xyz, lab = fut.result()
        if step + 1 < steps:
            fut = pool.submit(next_batch)
        if step < warmup:
            lr_t = lr * (step + 1) / warmup
        else:
            f = (step - warmup) / max(1, steps - warmup)
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * f))
        loss = prop.train_step(xyz, lab, lr_t)
        hist.append(loss)
        if (step + 1) % log_every == 0 or step + 1 == steps:
            log(f"step {step + 1}/{steps} loss {np.mean(hist[-log_every:]):.5f} lr {lr_t:.2e} "
                f"({time.time() - t0:.1f} s)")
    pool.shutdown()
    return prop.get_weights(), np.array(hist)


def evaluate(weights, ts: TrainingSet, n: int = 2048, seed: int = 99, device: int = 0, batch: int = 256):
    """Rotation error (deg) of the network's proposals on n augmented pairs."""
    prop = Proposer(max_pairs=batch, device=device, weights=weights)
    rng = np.random.default_rng(seed)
    errs = []
    for b0 in range(0, n, batch):
        idx = rng.integers(0, len(ts), size=min(batch, n - b0))
        xyz, lab = ts.batch(idx, rng)
        out = prop.forward_backward(xyz, None)
        errs.append(rotation_error_deg(out["R"], lab[:, :9].reshape(-1, 3, 3)))
    return np.concatenate(errs)


def _centroids(clouds):
    return np.stack([np.asarray(c, np.float64).mean(axis=0) for c in clouds])


class MultiStartRegistration:
    """Plain multi-start ICP: every pair from n_starts orientations spread
    over SO(3) about the centroids (multistart.so3_start_poses), best start
    by multistart.select_best."""

    def __init__(self, params=None, n_starts: int = 8, device: int = 0):
        self.params = params if params is not None else default_params()
        self.n_starts = n_starts
        self.reg = BatchRegistration(device)
        self.reg.set_params(self.params)

    def set_clouds(self, sources, targets, src_centroids=None, tgt_centroids=None):
        self.n = len(sources)
        self.reg.set_clouds(sources, targets)
        self.cs = _centroids(sources) if src_centroids is None else np.asarray(src_centroids, np.float64)
        self.ct = _centroids(targets) if tgt_centroids is None else np.asarray(tgt_centroids, np.float64)
        ids = np.repeat(np.arange(self.n), self.n_starts)
        self.reg.set_problems(ids, ids)
        self.guesses = np.stack([T for i in range(self.n)
                                 for T in multistart.so3_start_poses(self.cs[i], self.ct[i], self.n_starts)])

    def align(self):
        """Best start of every pair; self.inits holds the start pose each
        pair's result was registered from."""
        res = self.reg.align(self.guesses)
        self.iterations_total = int(sum(r["iterations"] for r in res))
        out = []
        best = np.empty(self.n, np.int64)
        for i in range(self.n):
            r = res[i * self.n_starts:(i + 1) * self.n_starts]
            best[i] = multistart.select_best(r)
            out.append(r[best[i]])
        self.inits = self.guesses[np.arange(self.n) * self.n_starts + best]
        return out


def default_params():
    """ICP settings of the registration workloads: max correspondence
    distance 0.05 m, <= 50 iterations, transformation epsilon 1e-8."""
    return make_params(max_iterations=50, max_correspondence_distance=0.05, transformation_epsilon=1e-8)


class LearnedRegistration:
    """Learned proposal -> ICP refine -> judge -> multi-start fallback: one
    call of the library's rtr_learned pipeline (C++, pipeline.cpp).

    With hypotheses = k > 1 the network runs on the source samples rotated by
    k fixed orientations spread over SO(3) (multistart.so3_rotations, the
    first the identity) and the candidate with the lowest truncated-chamfer
    score is the proposal ICP refines.

    refine_anderson > 0 runs the refine launch with Anderson acceleration of
    that memory (rtr_params.anderson_memory): an equally good minimum in
    fewer iterations; refine_coarse_stride > 1 starts it on every s-th source
    point for up to refine_coarse_iterations iterations (coarse-to-fine,
    rtr_params.coarse_stride); fallback_anderson does the same for the
    fallback's multi-start launch (0: the plain iteration).  The judge
    (rtr_batch_judge) rejects a pair whose plane RMSE is above rmse_threshold
    or whose inlier fraction is below min_inlier_frac; the rejected pairs run
    n_starts-start ICP (multistart.so3_rotations about the centroids) on an
    engine of their own, so the refine engine keeps one layout and replays
    its captured graph every call.  The refine engine's grids are built on a
    host thread while the network runs.

    `reg`, `fb` and `prop` are views of the pipeline's refine engine,
    fallback engine and network (for tuning and profiling)."""

    def __init__(self, weights=None, params=None, n_starts: int = 8, rmse_threshold: float = 2e-3,
                 min_inlier_frac: float = 0.8, judge_k: int = 20, judge_points: int = 1024,
                 max_curvature: float = 0.02, hypotheses: int = 8, tau: float = 0.1, max_pairs: int = 256,
                 refine_anderson: int = 5, fallback_item_pts: int = 256, fallback_blocks_per_cu: int = 1,
                 fallback_anderson: int = 0, refine_coarse_stride: int = 4, refine_coarse_iterations: int = 25,
                 device: int = 0):
        if weights is None:
            weights, _ = load_weights()
        self.params = L.Params.from_buffer_copy(params if params is not None else default_params())
        cfg = L.LearnedConfig()
        L.check(L.lib.rtr_learned_config_default(C.byref(cfg)))
        cfg.hypotheses, cfg.n_starts, cfg.tau = int(hypotheses), int(n_starts), float(tau)
        cfg.rmse_threshold, cfg.min_inlier_frac = float(rmse_threshold), float(min_inlier_frac)
        cfg.judge_k, cfg.judge_points, cfg.max_curvature = int(judge_k), int(judge_points), float(max_curvature)
        cfg.refine_anderson, cfg.refine_coarse_stride = int(refine_anderson), int(refine_coarse_stride)
        cfg.refine_coarse_iterations, cfg.fallback_anderson = int(refine_coarse_iterations), int(fallback_anderson)
        cfg.fallback_item_pts, cfg.fallback_blocks_per_cu = int(fallback_item_pts), int(fallback_blocks_per_cu)
        self.config = cfg
        h = C.c_void_p()
        L.check(L.lib.rtr_learned_create(C.byref(h), int(device), int(max_pairs), C.byref(self.params),
                                         C.byref(cfg)))
        self._h = h
        w = np.ascontiguousarray(weights, np.float32)
        if len(w) != n_params():
            raise ValueError(f"expected {n_params()} parameters, got {len(w)}")
        L.check(L.lib.rtr_learned_set_weights(h, L.ptr(w)))
        self.hypotheses = hypotheses
        self.n_starts = n_starts
        self.tau = tau
        self.rots = multistart.so3_rotations(hypotheses) if hypotheses > 1 else np.eye(3)[None]
        self.start_rots = multistart.so3_rotations(n_starts)
        hr = np.ascontiguousarray(self.rots, np.float64)
        sr = np.ascontiguousarray(self.start_rots, np.float64)
        L.check(L.lib.rtr_learned_set_rotations(h, L.ptr(hr), len(hr), L.ptr(sr), len(sr)))
        reg, fb, prop = C.c_void_p(), C.c_void_p(), C.c_void_p()
        L.check(L.lib.rtr_learned_batches(h, C.byref(reg), C.byref(fb), C.byref(prop)))
        self.reg = BatchRegistration.borrow(reg, self)
        self.fb = BatchRegistration.borrow(fb, self)
        self.prop = Proposer.borrow(prop, self, max_pairs * max(1, hypotheses))
        self._active = "refine"
        self.n = 0
        self.stats = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            L.lib.rtr_learned_destroy(h)
        self._h = None

    def set_clouds(self, sources, targets, src_centroids=None, tgt_centroids=None):
        src = [L.pts3(s) for s in sources]
        tgt = [L.pts3(t) for t in targets]
        if len(src) != len(tgt):
            raise ValueError("sources and targets must pair up")
        n = len(src)
        self.cs = _centroids(src) if src_centroids is None else np.ascontiguousarray(src_centroids, np.float64)
        self.ct = _centroids(tgt) if tgt_centroids is None else np.ascontiguousarray(tgt_centroids, np.float64)
        sp = (C.c_void_p * n)(*[a.ctypes.data for a in src])
        tp = (C.c_void_p * n)(*[a.ctypes.data for a in tgt])
        ns = (C.c_size_t * n)(*[len(a) for a in src])
        nt = (C.c_size_t * n)(*[len(a) for a in tgt])
        L.check(L.lib.rtr_learned_set_clouds(self._h, n, sp, ns, tp, nt, L.ptr(self.cs), L.ptr(self.ct)))
        self._clouds = (src, tgt)          # the library reads the rejected pairs' clouds from these
        self.n = n
        self.reg.n_problems = n
        self.reg._ns = [len(a) for a in src]
        self.reg._prob_src = np.arange(n, dtype=np.int32)
        self.reg._max_ids = (n - 1, n - 1)
        self.prop.n_pairs = n

    def use_params(self, which: str):
        """The refine engine runs the 'refine' launch or the 'plain' iteration
        (the same launch unaccelerated, the caller's ICP parameters)."""
        if which not in ("refine", "plain"):
            raise ValueError(which)
        L.check(L.lib.rtr_learned_use_plain_params(self._h, int(which == "plain")))
        self._active = which

    def _finish(self, res, inits, rej):
        st = L.LearnedStats()
        L.check(L.lib.rtr_learned_get_stats(self._h, C.byref(st)))
        n = self.n
        props = np.empty((n, 16))
        L.check(L.lib.rtr_learned_get_proposals(self._h, L.ptr(props)))
        rmse, frac, pfrac = np.empty(n), np.empty(n), np.empty(n)
        L.check(L.lib.rtr_learned_get_judge(self._h, L.ptr(rmse), L.ptr(frac), L.ptr(pfrac)))
        rowm = lambda a: np.ascontiguousarray(np.transpose(a.reshape(n, 4, 4), (0, 2, 1)))
        self.stats = {"n_fallback": int(st.n_fallback), "fallback_pairs": np.where(rej)[0],
                      "proposals": rowm(props), "inits": rowm(inits), "judge_rmse": rmse, "judge_frac": frac,
                      "judge_planar_frac": pfrac,
                      # ms_propose: the network, with the grid build running beside it;
                      # ms_build: what is left of the build after the network
                      "ms_build": st.ms_build_wait, "ms_propose": st.ms_propose, "ms_refine": st.ms_refine,
                      "ms_judge": st.ms_judge, "ms_fallback": st.ms_fallback, "ms_total": st.ms_total,
                      "ms_propose_device": st.ms_propose_device}
        self._active = "refine"
        return result_dicts(res)

    def align(self):
        """Register every pair: proposal, refine, judge, fallback."""
        n = self.n
        res = (L.Result * n)()
        inits = np.empty((n, 16))
        rej = np.zeros(n, np.int32)
        L.check(L.lib.rtr_learned_align(self._h, res, L.ptr(inits), L.ptr(rej)))
        return self._finish(res, inits, rej)

    def align_from(self, proposals):
        """Refine, judge and fall back from the caller's (n, 4, 4) proposals."""
        n = self.n
        g = np.ascontiguousarray(np.transpose(np.asarray(proposals, np.float64), (0, 2, 1))).reshape(-1)
        res = (L.Result * n)()
        inits = np.empty((n, 16))
        rej = np.zeros(n, np.int32)
        L.check(L.lib.rtr_learned_align_from(self._h, L.ptr(g), res, L.ptr(inits), L.ptr(rej)))
        return self._finish(res, inits, rej)
