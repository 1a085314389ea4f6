"""Synthetic workloads (numpy, seeded with PCG64(1000 * config + pair)).

The reference ships real scans only (RealTimeRobot/*.pcd); these generators
produce the registration workloads at the sizes the benchmarks use:
box-composite furniture sampled uniformly over its surface (area-weighted),
Gaussian noise, and random rigid motions.  All clouds are float32 (n, 3).
Transforms are 4x4 float64, row-major; ``T_gt`` maps the source onto the
target (the pose ICP should recover).
"""
from __future__ import annotations

import numpy as np


def rng_for(config: int, pair: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(1000 * config + pair))


# A box: (center xyz, size xyz, yaw radians)
CHAIR = [
    ((0.0, 0.0, 0.475), (0.45, 0.45, 0.05), 0.0),      # seat, top at 0.50
    ((0.205, 0.205, 0.225), (0.04, 0.04, 0.45), 0.0),   # legs
    ((-0.205, 0.205, 0.225), (0.04, 0.04, 0.45), 0.0),
    ((0.205, -0.205, 0.225), (0.04, 0.04, 0.45), 0.0),
    ((-0.205, -0.205, 0.225), (0.04, 0.04, 0.45), 0.0),
    ((0.0, 0.2, 0.75), (0.45, 0.05, 0.5), 0.0),        # backrest
]


def box_faces(boxes):
    """Per face: origin, two edge vectors and area."""
    faces = []
    for c, s, yaw in boxes:
        c = np.asarray(c, np.float64)
        s = np.asarray(s, np.float64)
        cy, sy = np.cos(yaw), np.sin(yaw)
        R = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1.0]])
        for ax in range(3):
            u, v = [a for a in range(3) if a != ax]
            for sign in (-1.0, 1.0):
                o = np.zeros(3)
                o[ax] = sign * s[ax] / 2
                o[u] = -s[u] / 2
                o[v] = -s[v] / 2
                eu = np.zeros(3)
                eu[u] = s[u]
                ev = np.zeros(3)
                ev[v] = s[v]
                faces.append((c + R @ o, R @ eu, R @ ev, s[u] * s[v]))
    return faces


def sample_boxes(boxes, n: int, rng: np.random.Generator, noise: float = 0.0) -> np.ndarray:
    faces = box_faces(boxes)
    areas = np.array([f[3] for f in faces])
    which = rng.choice(len(faces), size=n, p=areas / areas.sum())
    uv = rng.random((n, 2))
    O = np.stack([f[0] for f in faces])[which]
    U = np.stack([f[1] for f in faces])[which]
    V = np.stack([f[2] for f in faces])[which]
    p = O + U * uv[:, :1] + V * uv[:, 1:2]
    if noise > 0:
        p = p + rng.normal(0.0, noise, size=p.shape)
    return p.astype(np.float32)


def random_rotation(rng: np.random.Generator, max_deg: float) -> np.ndarray:
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    ang = np.deg2rad(rng.uniform(0.0, max_deg))
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


def random_rigid(rng: np.random.Generator, max_deg: float, max_t: float) -> np.ndarray:
    T = np.eye(4)
    T[:3, :3] = random_rotation(rng, max_deg)
    T[:3, 3] = rng.uniform(-max_t, max_t, size=3)
    return T


def apply(T, pts) -> np.ndarray:
    p = np.asarray(pts, np.float64)
    return (p @ T[:3, :3].T + T[:3, 3]).astype(np.float32)


def random_composite(rng: np.random.Generator, asymmetric: bool = False):
    """A random piece of box-composite furniture (0.4-2.0 m extents).  With
    asymmetric=True every table carries its off-centre drawer block, so no
    piece has a proper rotational symmetry (its pose is identifiable)."""
    kind = rng.integers(0, 3)
    ext = rng.uniform(0.4, 2.0)
    boxes = []
    if kind == 0:  # table/desk: top + legs
        w, d, h = ext, rng.uniform(0.4, 1.0) * ext, rng.uniform(0.4, 0.8)
        boxes.append(((0, 0, h - 0.02), (w, d, 0.04), 0.0))
        for sx in (-1, 1):
            for sy in (-1, 1):
                boxes.append(((sx * (w / 2 - 0.03), sy * (d / 2 - 0.03), (h - 0.04) / 2), (0.05, 0.05, h - 0.04), 0.0))
        if rng.random() < 0.5 or asymmetric:  # drawer block
            boxes.append(((w / 4, 0, h * 0.7), (w / 3, d * 0.9, h * 0.3), 0.0))
    elif kind == 1:  # chair with random proportions
        s = rng.uniform(0.4, 0.6)
        h = rng.uniform(0.4, 0.5)
        boxes.append(((0, 0, h), (s, s, 0.05), 0.0))
        for sx in (-1, 1):
            for sy in (-1, 1):
                boxes.append(((sx * (s / 2 - 0.03), sy * (s / 2 - 0.03), h / 2), (0.04, 0.04, h), 0.0))
        bh = rng.uniform(0.3, 0.6)
        boxes.append(((0, s / 2 - 0.025, h + bh / 2), (s, 0.05, bh), 0.0))
    else:  # sofa / cabinet: base + back + arms
        w, d = ext, rng.uniform(0.5, 0.9)
        boxes.append(((0, 0, 0.22), (w, d, 0.44), 0.0))
        boxes.append(((0, d / 2 - 0.1, 0.6), (w, 0.2, 0.5), 0.0))
        if rng.random() < 0.7:
            boxes.append(((-w / 2 + 0.08, 0, 0.5), (0.16, d, 0.3), 0.0))
            boxes.append(((w / 2 - 0.08, 0, 0.5), (0.16, d, 0.3), 0.0))
    return boxes


def loguniform_int(rng, lo, hi):
    return int(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def config1_pair(pair: int = 0):
    """Chair target 20,000 pts + N(0, 2 mm); source an independent 10,000-pt
    resample moved by T_src (axis on S2, angle U[0, 15 deg], t U[-0.1, 0.1]).
    Returns (source, target, T_gt) with T_gt = inverse(T_src)."""
    rng = rng_for(1, pair)
    tgt = sample_boxes(CHAIR, 20000, rng, noise=0.002)
    src0 = sample_boxes(CHAIR, 10000, rng, noise=0.002)
    T_src = random_rigid(rng, 15.0, 0.1)
    return apply(T_src, src0), tgt, np.linalg.inv(T_src)


def config3_batch(n_pairs: int = 256, seed_offset: int = 0, ns_range=(2000, 50000), nt_range=(10000, 200000),
                  max_deg: float = 30.0, max_t: float = 0.1):
    """Independent pairs of random box composites with log-uniform sizes;
    the source is moved by a random rotation U[0, max_deg] about a random
    axis and a translation U[-max_t, max_t]^3.  Returns lists (sources,
    targets, T_gt)."""
    srcs, tgts, gts = [], [], []
    for p in range(n_pairs):
        rng = rng_for(3, seed_offset + p)
        boxes = random_composite(rng)
        ns = loguniform_int(rng, *ns_range)
        nt = loguniform_int(rng, *nt_range)
        tgt = sample_boxes(boxes, nt, rng, noise=0.002)
        src0 = sample_boxes(boxes, ns, rng, noise=0.002)
        T_src = random_rigid(rng, max_deg, max_t)
        srcs.append(apply(T_src, src0))
        tgts.append(tgt)
        gts.append(np.linalg.inv(T_src))
    return srcs, tgts, gts


def rmse_between(T_a, T_b, pts) -> float:
    """sqrt(mean ||T_a p - T_b p||^2) over the points (FP64)."""
    p = np.asarray(pts, np.float64)
    d = p @ (T_a[:3, :3] - T_b[:3, :3]).T + (T_a[:3, 3] - T_b[:3, 3])
    return float(np.sqrt(np.mean(np.sum(d * d, axis=1))))


_FACE_AX = np.array([0, 0, 1, 1, 2, 2])
_FACE_SIGN = np.array([-1.0, 1.0, -1.0, 1.0, -1.0, 1.0])


def _faces_array(boxes):
    """Vectorised box_faces: origins, edge vectors (n_faces, 3) and areas."""
    c = np.array([b[0] for b in boxes], np.float64)
    sz = np.array([b[1] for b in boxes], np.float64)
    yaw = np.array([b[2] for b in boxes], np.float64)
    nb = len(boxes)
    cy, sy = np.cos(yaw), np.sin(yaw)
    R = np.zeros((nb, 3, 3))
    R[:, 0, 0], R[:, 0, 1], R[:, 1, 0], R[:, 1, 1], R[:, 2, 2] = cy, -sy, sy, cy, 1.0
    ax = np.tile(_FACE_AX, nb)
    sign = np.tile(_FACE_SIGN, nb)
    bi = np.repeat(np.arange(nb), 6)
    u = (ax + 1) % 3
    v = (ax + 2) % 3
    u, v = np.minimum(u, v), np.maximum(u, v)
    s6 = sz[bi]
    k = np.arange(6 * nb)
    o = np.zeros((6 * nb, 3))
    o[k, ax] = sign * s6[k, ax] / 2
    o[k, u] = -s6[k, u] / 2
    o[k, v] = -s6[k, v] / 2
    eu = np.zeros((6 * nb, 3))
    eu[k, u] = s6[k, u]
    ev = np.zeros((6 * nb, 3))
    ev[k, v] = s6[k, v]
    Rb = R[bi]
    O = c[bi] + np.einsum("kij,kj->ki", Rb, o)
    return O, np.einsum("kij,kj->ki", Rb, eu), np.einsum("kij,kj->ki", Rb, ev), s6[k, u] * s6[k, v]


def sample_many(pieces, counts, rng: np.random.Generator, noise: float = 0.0):
    """Area-weighted surface samples of many box composites at once (the
    points of a cloud come grouped by face).  Returns one float32
    (counts[i], 3) array per piece (piece frame)."""
    counts = np.asarray(counts, np.int64)
    Os, Us, Vs, per_face = [], [], [], []
    cache = {}
    for c, boxes in enumerate(pieces):
        key = id(boxes)
        if key not in cache:
            cache[key] = _faces_array(boxes)
        O, U, V, a = cache[key]
        Os.append(O)
        Us.append(U)
        Vs.append(V)
        per_face.append(rng.multinomial(int(counts[c]), a / a.sum()))
    O = np.concatenate(Os).astype(np.float32)
    U = np.concatenate(Us).astype(np.float32)
    V = np.concatenate(Vs).astype(np.float32)
    nf = np.concatenate(per_face)
    uv = rng.random((int(nf.sum()), 2), dtype=np.float32)
    p = np.repeat(O, nf, axis=0)
    p += np.repeat(U, nf, axis=0) * uv[:, :1]
    p += np.repeat(V, nf, axis=0) * uv[:, 1:2]
    if noise > 0:
        p += np.float32(noise) * rng.standard_normal(p.shape, dtype=np.float32)
    return np.split(p, np.cumsum(counts)[:-1])


def config4_batch(n_pairs: int = 256, seed_offset: int = 0, ns_range=(2000, 50000), nt_range=(10000, 200000),
                  max_deg: float = 180.0, max_t: float = 0.1, noise: float = 0.001):
    """Learned-proposal workload: pairs of asymmetric random box composites
    with log-uniform sizes (as the batched workload), N(0, noise) per axis,
    the source moved by a rotation U[0, max_deg] about a uniform axis and a
    translation U[-max_t, max_t]^3.  Returns (sources, targets, T_gt)."""
    srcs, tgts, gts = [], [], []
    for p in range(n_pairs):
        rng = rng_for(4, seed_offset + p)
        boxes = random_composite(rng, asymmetric=True)
        ns = loguniform_int(rng, *ns_range)
        nt = loguniform_int(rng, *nt_range)
        tgt, src0 = sample_many([boxes, boxes], [nt, ns], rng, noise)
        T_src = random_rigid(rng, max_deg, max_t)
        srcs.append(apply(T_src, src0))
        tgts.append(tgt)
        gts.append(np.linalg.inv(T_src))
    return srcs, tgts, gts


def training_base(n_pairs: int, seed: int, ns_range=(2000, 6000), nt_range=(4000, 12000), noise: float = 0.001,
                  chunk: int = 512):
    """Canonical training pairs (source and target independent samples of the
    same asymmetric composite, both in the piece frame), yielded in chunks of
    (sources, targets)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    done = 0
    while done < n_pairs:
        m = min(chunk, n_pairs - done)
        pieces = [random_composite(rng, asymmetric=True) for _ in range(m)]
        ns = [loguniform_int(rng, *ns_range) for _ in range(m)]
        nt = [loguniform_int(rng, *nt_range) for _ in range(m)]
        clouds = sample_many(pieces + pieces, nt + ns, rng, noise)
        yield clouds[m:], clouds[:m]
        done += m
