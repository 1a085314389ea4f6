"""CPU restatement of the learned pose proposer's network.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and the smoke check) as the
checker of the HIP kernels in realtime_robot_amd/csrc/proposer.hip; never
called by the product path.

The proposer has no reference counterpart (it replaces the reference's
keypoint/RANSAC coarse alignment, RealTimeRobot/function.h:35-109), so this
module restates the network from its definition:

  per cloud (1024 centred points, xyz in channels 0-2 of 16):
      Y1 = relu(X0 W1 + b1)  (64)   Y2 = relu(Y1 W2 + b2)  (128)
      Y3 = relu(Y2 W3 + b3)  (256)  G = max over the 1024 rows (argmax: lowest row)
  per pair: F = [G_src, G_tgt] (512);  H1 = relu(F W4 + b4) (256);  O = H1 W5 + b5 (9)
  head: R = Gram-Schmidt(O[0:3], O[3:6]) (columns b1, b2, b1 x b2), t = O[6:9]
  loss (mean over pairs) = mean_i ||(R - R*) s_i + (t - t*)||^2 over the source points s_i.

``emulate_bf16=True`` rounds to bfloat16 (round-to-nearest-even) exactly where
the device stores BF16: inputs, weights, the activations after bias+ReLU, the
pooled features, and the gradients fed to products.  Accumulation is FP64
here (FP32 on the device).  With ``emulate_bf16=False`` the function is the
exact FP64 network whose analytic gradient the tests pin by central finite
differences.
"""
from __future__ import annotations

import numpy as np

KP, C0, C1, C2, C3, H, OUT = 1024, 16, 64, 128, 256, 256, 9

_SHAPES = [("W1", (C0, C1)), ("b1", (C1,)), ("W2", (C1, C2)), ("b2", (C2,)), ("W3", (C2, C3)), ("b3", (C3,)),
           ("W4", (2 * C3, H)), ("b4", (H,)), ("W5", (H, OUT)), ("b5", (OUT,))]


def n_params() -> int:
    return int(sum(np.prod(s) for _, s in _SHAPES))


def unpack(w):
    """Views of the flat parameter vector (layout of rtr_prop_num_params)."""
    out, o = {}, 0
    for name, shp in _SHAPES:
        n = int(np.prod(shp))
        out[name] = w[o:o + n].reshape(shp)
        o += n
    return out


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(x):
    return np.asarray(x, np.float64)


def gram_schmidt(o):
    """(n, 9) head outputs -> R (n, 3, 3) with columns b1, b2, b3 and t (n, 3)."""
    a1, a2 = o[:, 0:3], o[:, 3:6]
    n1 = np.linalg.norm(a1, axis=1, keepdims=True) + 1e-12
    b1 = a1 / n1
    c = np.sum(b1 * a2, axis=1, keepdims=True)
    u2 = a2 - c * b1
    n2 = np.linalg.norm(u2, axis=1, keepdims=True) + 1e-12
    b2 = u2 / n2
    b3 = np.cross(b1, b2)
    R = np.stack([b1, b2, b3], axis=2)
    return R, o[:, 6:9].copy()


def head_loss(o, src_pts, labels):
    """Loss per pair and dL/dO of the mean loss.  src_pts (n, KP, 3); labels
    (n, 12): R* row-major then t*."""
    n = o.shape[0]
    R, t = gram_schmidt(o)
    Rs = labels[:, :9].reshape(n, 3, 3)
    ts = labels[:, 9:12]
    S = np.einsum("nia,nib->nab", src_pts, src_pts) / src_pts.shape[1]
    m = src_pts.mean(axis=1)
    dR = R - Rs
    dt = t - ts
    dRm = np.einsum("nra,na->nr", dR, m)
    loss = np.einsum("nra,nab,nrb->n", dR, S, dR) + 2 * np.sum(dt * dRm, axis=1) + np.sum(dt * dt, axis=1)
    gR = (2 * np.einsum("nrb,nba->nra", dR, S) + 2 * dt[:, :, None] * m[:, None, :]) / n
    gt = (2 * dt + 2 * dRm) / n
    # back through Gram-Schmidt
    a1, a2 = o[:, 0:3], o[:, 3:6]
    n1 = np.linalg.norm(a1, axis=1, keepdims=True) + 1e-12
    b1 = a1 / n1
    c = np.sum(b1 * a2, axis=1, keepdims=True)
    u2 = a2 - c * b1
    n2 = np.linalg.norm(u2, axis=1, keepdims=True) + 1e-12
    b2 = u2 / n2
    g1, g2, g3 = gR[:, :, 0].copy(), gR[:, :, 1].copy(), gR[:, :, 2]
    g1 += np.cross(b2, g3)
    g2 += np.cross(g3, b1)
    dU = (g2 - b2 * np.sum(b2 * g2, axis=1, keepdims=True)) / n2
    dUb1 = np.sum(dU * b1, axis=1, keepdims=True)
    da2 = dU - dUb1 * b1
    g1 += -dUb1 * a2 - c * dU
    da1 = (g1 - b1 * np.sum(b1 * g1, axis=1, keepdims=True)) / n1
    dO = np.concatenate([da1, da2, gt], axis=1)
    return loss, dO, R, t


def forward_backward(w, xyz, labels=None, emulate_bf16=True):
    """xyz: (n_pairs * 2 * KP, 3) centred FPS points, clouds ordered src0, tgt0,
    src1, ...  Returns dict with O, R, t and (with labels) loss (mean) and
    grad (flat, same layout as w)."""
    q = bf16 if emulate_bf16 else _id
    P = unpack(np.asarray(w, np.float64))
    Wb = {k: q(v) if k.startswith("W") else np.asarray(v, np.float64) for k, v in P.items()}
    if emulate_bf16:  # biases are FP32 on the device
        Wb.update({k: np.float32(v).astype(np.float64) for k, v in P.items() if k.startswith("b")})
    xyz = np.asarray(xyz, np.float32)
    rows = xyz.shape[0]
    n_clouds = rows // KP
    n_pairs = n_clouds // 2
    X0 = np.zeros((rows, C0))
    X0[:, :3] = xyz
    X0 = q(X0)
    A1 = X0 @ Wb["W1"] + Wb["b1"]
    Y1 = q(np.maximum(A1, 0))
    A2 = Y1 @ Wb["W2"] + Wb["b2"]
    Y2 = q(np.maximum(A2, 0))
    A3 = Y2 @ Wb["W3"] + Wb["b3"]
    Y3 = q(np.maximum(A3, 0))
    Y3c = Y3.reshape(n_clouds, KP, C3)
    arg = np.argmax(Y3c, axis=1)                     # first (lowest) row on ties
    G = np.take_along_axis(Y3c, arg[:, None, :], axis=1)[:, 0, :]
    F = q(G.reshape(n_pairs, 2 * C3))
    A4 = F @ Wb["W4"] + Wb["b4"]
    H1 = q(np.maximum(A4, 0))
    O = H1 @ Wb["W5"] + Wb["b5"]
    if emulate_bf16:
        O = O.astype(np.float32).astype(np.float64)
    out = {"O": O}
    src_pts = xyz.reshape(n_pairs, 2, KP, 3)[:, 0].astype(np.float64)
    if labels is None:
        R, t = gram_schmidt(O)
        out.update(R=R, t=t)
        return out
    labels = np.asarray(labels, np.float64).reshape(n_pairs, 12)
    loss, dO, R, t = head_loss(O, src_pts, labels)
    out.update(R=R, t=t, loss_per_pair=loss, loss=float(loss.mean()))
    if emulate_bf16:
        dO = dO.astype(np.float32).astype(np.float64)
    g = {}
    dOb = q(dO)
    g["W5"] = H1.T @ dOb
    g["b5"] = dO.sum(axis=0)
    dH1 = dOb @ Wb["W5"].T
    dA4 = q(dH1 * (H1 > 0))
    g["W4"] = F.T @ dA4
    g["b4"] = dA4.sum(axis=0)
    dF = dA4 @ Wb["W4"].T
    dG = dF.reshape(n_clouds, C3)
    dA3 = np.zeros((n_clouds, KP, C3))
    ci, ch = np.meshgrid(np.arange(n_clouds), np.arange(C3), indexing="ij")
    dA3[ci, arg, ch] = np.where(G > 0, dG, 0.0)
    dA3 = q(dA3.reshape(rows, C3))
    g["W3"] = Y2.T @ dA3
    g["b3"] = dA3.sum(axis=0)
    dY2 = dA3 @ Wb["W3"].T
    dA2 = q(dY2 * (Y2 > 0))
    g["W2"] = Y1.T @ dA2
    g["b2"] = dA2.sum(axis=0)
    dY1 = dA2 @ Wb["W2"].T
    dA1 = q(dY1 * (Y1 > 0))
    g["W1"] = X0.T @ dA1
    g["b1"] = dA1.sum(axis=0)
    out["grad"] = np.concatenate([g[name].reshape(-1) for name, _ in _SHAPES])
    return out

