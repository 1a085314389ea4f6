"""CPU checks of the learned proposer's oracle and host logic (no GPU):
the oracle network's analytic gradient against central finite differences,
its BF16 rounding against torch's, the label composition used for training
augmentation, the multi-start orientation set and the committed weights."""
import os

import numpy as np
import pytest

from oracle import mlp
from realtime_robot_amd import learned, multistart


def _random_problem(n_pairs=1, seed=0):
    rng = np.random.default_rng(seed)
    w = np.concatenate([rng.normal(0, 0.3, size=int(np.prod(s))) for _, s in mlp._SHAPES])
    xyz = rng.normal(0, 0.3, size=(n_pairs * 2 * mlp.KP, 3)).astype(np.float32)
    lab = np.zeros((n_pairs, 12))
    for i in range(n_pairs):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        q *= np.sign(np.linalg.det(q))
        lab[i, :9] = q.reshape(-1)
        lab[i, 9:] = rng.normal(0, 0.1, 3)
    return w, xyz, lab, rng


def test_oracle_gradient_matches_finite_differences():
    w, xyz, lab, rng = _random_problem(2, 0)
    g = mlp.forward_backward(w, xyz, lab, emulate_bf16=False)["grad"]
    for _ in range(3):
        d = rng.normal(size=w.shape)
        eps = 1e-6
        lp = mlp.forward_backward(w + eps * d, xyz, lab, False)["loss"]
        lm = mlp.forward_backward(w - eps * d, xyz, lab, False)["loss"]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-6 * abs(fd)
    # every parameter block on its own (a transposed or misrouted block fails here)
    off = 0
    for name, shp in mlp._SHAPES:
        sz = int(np.prod(shp))
        d = np.zeros_like(w)
        d[off:off + sz] = rng.normal(size=sz)
        if name == "W1":
            d[off + 3 * 64:off + sz] = 0.0      # rows 3-15 see zero inputs
        eps = 1e-6
        fd = (mlp.forward_backward(w + eps * d, xyz, lab, False)["loss"] -
              mlp.forward_backward(w - eps * d, xyz, lab, False)["loss"]) / (2 * eps)
        an = g @ d
        assert abs(fd - an) <= 1e-5 * max(abs(fd), 1e-3), name
        off += sz


def test_head_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    o = rng.normal(size=(4, 9))
    pts = rng.normal(0, 0.4, size=(4, 1024, 3))
    lab = np.concatenate([np.stack([np.linalg.qr(rng.normal(size=(3, 3)))[0].reshape(-1) for _ in range(4)]),
                          rng.normal(0, 0.1, (4, 3))], axis=1)
    loss, dO, _, _ = mlp.head_loss(o, pts, lab)
    for i in range(4):
        for j in range(9):
            e = np.zeros_like(o)
            e[i, j] = 1e-6
            fd = (mlp.head_loss(o + e, pts, lab)[0].mean() - mlp.head_loss(o - e, pts, lab)[0].mean()) / 2e-6
            assert abs(fd - dO[i, j]) <= 1e-6 * max(1.0, abs(fd))


def test_head_recovers_rotation_for_exact_6d():
    rng = np.random.default_rng(4)
    Rs = np.stack([np.linalg.qr(rng.normal(size=(3, 3)))[0] for _ in range(5)])
    Rs *= np.sign(np.linalg.det(Rs))[:, None, None]
    o = np.concatenate([Rs[:, :, 0] * 2.0, Rs[:, :, 1] + 0.3 * Rs[:, :, 0], rng.normal(size=(5, 3))], axis=1)
    R, t = mlp.gram_schmidt(o)
    assert np.abs(R - Rs).max() < 1e-11 and np.array_equal(t, o[:, 6:])


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(size=10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)
    assert mlp.bf16(np.float32(1.0 + 2 ** -8)) == 1.0                 # tie -> even
    assert mlp.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def test_augmented_labels_map_source_onto_target():
    """Base pair with identical centred clouds and the identity label: after
    a random rigid motion A of the source, the composed label must map every
    moved source point back onto its target point."""
    rng = np.random.default_rng(6)
    P = rng.normal(0, 0.3, size=(4, 1024, 3)).astype(np.float32)
    c = rng.normal(0, 0.5, size=(4, 3))
    ts = learned.TrainingSet(P, P.copy(), c, c.copy(), np.tile(np.eye(4), (4, 1, 1)))
    xyz, lab = ts.batch(np.arange(4), rng)
    X = xyz.reshape(4, 2, 1024, 3).astype(np.float64)
    for i in range(4):
        Rl = lab[i, :9].reshape(3, 3)
        moved = X[i, 0] @ Rl.T + lab[i, 9:]
        assert np.abs(moved - X[i, 1]).max() < 1e-5
        assert abs(np.linalg.det(Rl) - 1) < 1e-9


def test_centred_labels_round_trip():
    rng = np.random.default_rng(7)
    T = learned.random_rigid_batch(rng, 6, 180.0, 0.5)
    cs, ct = rng.normal(size=(6, 3)), rng.normal(size=(6, 3))
    lab = learned.centred_labels(T, cs, ct)
    for i in range(6):
        Rl, tl = lab[i, :9].reshape(3, 3), lab[i, 9:]
        # T = Tr(ct) [R | t*] Tr(-cs)
        rec = np.eye(4)
        rec[:3, :3] = Rl
        rec[:3, 3] = tl + ct[i] - Rl @ cs[i]
        assert np.abs(rec - T[i]).max() < 1e-12


def test_random_rigid_batch_angles_cover_range():
    rng = np.random.default_rng(8)
    T = learned.random_rigid_batch(rng, 4000, 180.0, 0.1)
    ang = np.degrees(np.arccos(np.clip((np.trace(T[:, :3, :3], axis1=1, axis2=2) - 1) / 2, -1, 1)))
    assert ang.min() >= 0 and ang.max() <= 180.0 + 1e-9
    assert abs(np.mean(ang) - 90.0) < 3.0          # U[0, 180]
    assert np.abs(T[:, :3, 3]).max() <= 0.1


def test_so3_starts_are_spread_and_proper():
    R = multistart.so3_rotations(8)
    assert np.allclose(R[0], np.eye(3))
    assert np.allclose(np.einsum("kji,kjl->kil", R, R), np.eye(3), atol=1e-12)
    assert np.allclose(np.linalg.det(R), 1.0)
    ang = np.degrees(np.arccos(np.clip((np.einsum("aij,bij->ab", R, R) - 1) / 2, -1, 1)))
    off = ang[~np.eye(8, dtype=bool)]
    assert off.min() > 100.0
    poses = multistart.so3_start_poses([1.0, 2.0, 3.0], [0.5, -1.0, 0.2])
    for T in poses:   # source centroid lands on the target centroid
        assert np.allclose(T[:3, :3] @ [1.0, 2.0, 3.0] + T[:3, 3], [0.5, -1.0, 0.2])


def test_init_weights_layout():
    w = learned.init_weights(0)
    assert len(w) == mlp.n_params() == learned.n_params()
    P = mlp.unpack(w.astype(np.float64))
    assert np.array_equal(P["b5"], [1, 0, 0, 0, 1, 0, 0, 0, 0])
    assert all(np.all(P[k] == 0) for k in ("b1", "b2", "b3", "b4"))


@pytest.mark.skipif(not os.path.exists(learned.WEIGHTS_PATH), reason="no trained weights in the tree")
def test_committed_weights_load():
    w, meta = learned.load_weights()
    assert len(w) == learned.n_params() and np.all(np.isfinite(w))
    assert meta["pairs_seen"] >= 200000


def test_fps_oracle_is_greedy_farthest_point():
    """Pins oracle.fps by what defines farthest-point sampling: every pick
    maximises the distance to the points already chosen (checked in FP64 by
    brute force), picks are distinct while points remain, and the covering
    radius never grows."""
    import oracle
    rng = np.random.default_rng(9)
    pts = rng.normal(0, 0.5, size=(1500, 3)).astype(np.float32)
    sel, cen, sub = oracle.fps(pts, k=300, max_sub=1200)
    assert len(sub) == 1200 and sel[0] == 0
    assert np.allclose(cen, sub.astype(np.float64).mean(0), atol=1e-12)
    assert np.array_equal(sub, pts[(np.arange(1200) * 1500) // 1200])
    assert len(np.unique(sel)) == 300
    S = sub.astype(np.float64)
    d = np.sum((S - S[sel[0]]) ** 2, axis=1)
    radii = []
    for s in range(1, 300):
        best = d.max()
        radii.append(best)
        assert d[sel[s]] >= best * (1 - 1e-5)
        d = np.minimum(d, np.sum((S - S[sel[s]]) ** 2, axis=1))
    assert np.all(np.diff(radii) <= 1e-12)
    # fewer points than picks: every point first, then index 0 repeats
    sel2, _, _ = oracle.fps(pts[:50], k=64)
    assert sorted(sel2[:50]) == list(range(50)) and np.all(sel2[50:] == 0)
