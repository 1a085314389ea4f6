"""Synthetic workloads (numpy, seeded with PCG64(1000 * config + pair)).

The reference ships real scans only (RealTimeRobot/*.pcd); these generators
produce the registration workloads at the sizes the benchmarks use:
box-composite furniture sampled uniformly over its surface (area-weighted),
Gaussian noise, and random rigid motions.  All clouds are float32 (n, 3).
Transforms are 4x4 float64, row-major; ``T_gt`` maps the source onto the
target (the pose ICP should recover).
"""
from __future__ import annotations

import os

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


def config4_yaw_batch(n_pairs: int = 256, seed_offset: int = 0, ns_range=(2000, 50000), nt_range=(10000, 200000),
                      max_deg: float = 180.0, max_t: float = 0.1, noise: float = 0.001):
    """config4_batch with the source turned about the vertical (z) axis only,
    by U[-max_deg, max_deg]: the gravity-aligned setting the reference's own
    coarse alignment (a yaw search) assumes.  Returns (sources, targets, T_gt)."""
    srcs, tgts, gts = [], [], []
    for p in range(n_pairs):
        rng = rng_for(41, seed_offset + p)
        boxes = random_composite(rng, asymmetric=True)
        ns = loguniform_int(rng, *ns_range)
        nt = loguniform_int(rng, *nt_range)
        tgt, src0 = sample_many([boxes, boxes], [nt, ns], rng, noise)
        a = np.deg2rad(rng.uniform(-max_deg, max_deg))
        T_src = np.eye(4)
        T_src[:3, :3] = [[np.cos(a), -np.sin(a), 0.0], [np.sin(a), np.cos(a), 0.0], [0.0, 0.0, 1.0]]
        T_src[:3, 3] = rng.uniform(-max_t, max_t, size=3)
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


def room_faces(size=(5.0, 5.0, 2.5)):
    """Floor (z = 0) and four walls of a room [0, sx] x [0, sy] x [0, sz] as
    (origin, edge u, edge v, area) faces."""
    sx, sy, sz = size
    o = np.zeros(3)
    return [(o, np.array([sx, 0, 0.0]), np.array([0, sy, 0.0]), sx * sy),
            (o, np.array([sx, 0, 0.0]), np.array([0, 0, sz]), sx * sz),
            (np.array([0, sy, 0.0]), np.array([sx, 0, 0.0]), np.array([0, 0, sz]), sx * sz),
            (o, np.array([0, sy, 0.0]), np.array([0, 0, sz]), sy * sz),
            (np.array([sx, 0, 0.0]), np.array([0, sy, 0.0]), np.array([0, 0, sz]), sy * sz)]


def sample_faces(faces, n, rng, noise=0.0):
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


def floor_pose(rng, size=(5.0, 5.0, 2.5), margin=1.0):
    """A furniture pose on the floor: yaw U[0, 360 deg], position inside the room."""
    T = np.eye(4)
    yaw = rng.uniform(0, 2 * np.pi)
    c, s = np.cos(yaw), np.sin(yaw)
    T[:2, :2] = [[c, -s], [s, c]]
    T[0, 3] = rng.uniform(margin, size[0] - margin)
    T[1, 3] = rng.uniform(margin, size[1] - margin)
    return T


def config2_scene(seed: int = 0, n_room: int = 300000, n_model: int = 30000, density: float = 4000.0,
                  noise: float = 0.003, max_deg: float = 20.0, max_t: float = 0.2):
    """Model-to-scene point-to-plane workload: a 5 x 5 x 2.5 m room (floor +
    4 walls) sampled at n_room points plus 6 box-composite pieces standing on
    the floor (the first is the chair that the model is, the others random
    furniture, 0.4-2.0 m) sampled at the same surface density, noise
    N(0, noise) per axis; the model is an independent n_model-point sample
    of the chair in its own frame.  Returns (model, scene, T_gt, T_init):
    T_gt maps the model onto its instance in the scene, T_init is T_gt
    perturbed by a rotation U[0, max_deg] about the model centroid and a
    translation U[-max_t, max_t]^3."""
    rng = rng_for(2, seed)
    parts = [sample_faces(room_faces(), n_room, rng, noise)]
    T_gt = floor_pose(rng)
    poses = [T_gt] + [floor_pose(rng) for _ in range(5)]
    pieces = [CHAIR] + [random_composite(rng) for _ in range(5)]
    for T, boxes in zip(poses, pieces):
        area = sum(f[3] for f in box_faces(boxes))
        pts = sample_boxes(boxes, int(area * density), rng, noise)
        parts.append(apply(T, pts))
    scene = np.concatenate(parts)
    model = sample_boxes(CHAIR, n_model, rng, 0.001)
    D = random_rigid(rng, max_deg, 0.0)
    c = model.astype(np.float64).mean(0)
    Dm = np.eye(4)
    Dm[:3, :3] = D[:3, :3]
    Dm[:3, 3] = c - D[:3, :3] @ c + rng.uniform(-max_t, max_t, 3)
    return model, scene, T_gt, T_gt @ Dm


# ---------------------------------------------------------------------------
# depth rendering of the room scene (streaming reconstruction workload)
# ---------------------------------------------------------------------------

def room_boxes(seed: int = 0):
    """The config-2 room's furniture as world-frame oriented boxes
    (centre, half extents, 3x3 rotation), same random draws as config2_scene."""
    rng = rng_for(2, seed)
    sample_faces(room_faces(), 1, rng, 0.0)  # keep the generator's draw order
    T_gt = floor_pose(rng)
    poses = [T_gt] + [floor_pose(rng) for _ in range(5)]
    pieces = [CHAIR] + [random_composite(rng) for _ in range(5)]
    out = []
    for T, boxes in zip(poses, pieces):
        for c, s, yaw in boxes:
            cy, sy = np.cos(yaw), np.sin(yaw)
            Rb = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1.0]])
            out.append((T[:3, :3] @ np.asarray(c, np.float64) + T[:3, 3], np.asarray(s, np.float64) / 2,
                        T[:3, :3] @ Rb))
    return out


def camera_path(n_frames: int, seed: int = 0, size=(5.0, 5.0, 2.5), step: float = 0.9 * np.pi / 63):
    """A smooth 6-DoF camera trajectory inside the room: the camera moves on
    an ellipse at 1.2-1.6 m height and looks towards the room centre with a
    gentle roll/pitch wobble, advancing `step` radians along the ellipse per
    frame (64 frames cover 0.9 pi; shorter runs are prefixes of that path).
    Returns camera-to-world poses (n, 4, 4); the camera looks along +z, x
    right, y down."""
    rng = rng_for(5, seed)
    ph = rng.uniform(0, 2 * np.pi)
    poses = []
    for k in range(n_frames):
        a = ph + step * k
        pos = np.array([size[0] / 2 + 1.3 * np.cos(a), size[1] / 2 + 1.1 * np.sin(a),
                        1.4 + 0.2 * np.sin(3.1 * a)])
        target = np.array([size[0] / 2 + 0.6 * np.cos(a + 1.0), size[1] / 2 + 0.6 * np.sin(0.8 * a), 0.6])
        z = target - pos
        z /= np.linalg.norm(z)
        up = np.array([0, 0, 1.0])
        x = np.cross(z, up)
        x /= np.linalg.norm(x)
        y = np.cross(z, x)
        R = np.stack([x, y, z], axis=1)
        roll = 0.08 * np.sin(2.3 * a)
        cr, sr = np.cos(roll), np.sin(roll)
        R = R @ np.array([[cr, -sr, 0], [sr, cr, 0], [0, 0, 1.0]])
        T = np.eye(4)
        T[:3, :3] = R
        T[:3, 3] = pos
        poses.append(T)
    return np.stack(poses)


def render_depth(T_cw, boxes, width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5,
                 size=(5.0, 5.0, 2.5), noise_rng=None, noise_k: float = 0.0015, max_depth: float = 8.0):
    """Depth image (metres along the optical axis, 0 = no hit) of the room
    (floor + 4 walls, open ceiling) and the boxes seen from camera-to-world
    pose T_cw, with depth noise N(0, noise_k * z^2) when noise_rng is given."""
    z = render_depth_f64(T_cw, boxes, width, height, fx, fy, cx, cy, size, max_depth)
    if noise_rng is not None:
        add_depth_noise(z, noise_rng, noise_k)
    return z.reshape(height, width).astype(np.float32)


def _rows_times(A, M):
    """A @ M for (n, 3) A and 3x3 M, summed left to right elementwise."""
    return A[:, 0:1] * M[0] + A[:, 1:2] * M[1] + A[:, 2:3] * M[2]


def render_depth_f64(T_cw, boxes, width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5,
                     size=(5.0, 5.0, 2.5), max_depth: float = 8.0):
    """The noise-free depth of render_depth as a flat float64 array."""
    u, v = np.meshgrid(np.arange(width, dtype=np.float64), np.arange(height, dtype=np.float64))
    d_cam = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)], axis=-1).reshape(-1, 3)
    R, o = T_cw[:3, :3], T_cw[:3, 3]
    # explicit products (no BLAS): the same bits on every machine and safe to
    # call from several threads at once
    d = _rows_times(d_cam, R.T)
    t_best = np.full(len(d), np.inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        # room: floor z = 0 and the four walls, hit from inside
        for ax, val in ((2, 0.0), (0, 0.0), (0, size[0]), (1, 0.0), (1, size[1])):
            t = (val - o[ax]) / d[:, ax]
            p = o + t[:, None] * d
            ok = (t > 1e-6) & np.all((p >= -1e-9) & (p <= np.array(size) + 1e-9), axis=1)
            t_best = np.where(ok & (t < t_best), t, t_best)
        for c, h, Rb in boxes:
            oc = o - c
            ol = oc[0] * Rb[0] + oc[1] * Rb[1] + oc[2] * Rb[2]
            dl = _rows_times(d, Rb)
            t1 = (-h - ol) / dl
            t2 = (h - ol) / dl
            tmin = np.nanmax(np.minimum(t1, t2), axis=1)
            tmax = np.nanmin(np.maximum(t1, t2), axis=1)
            ok = (tmax >= np.maximum(tmin, 0.0)) & (tmin > 1e-6)
            t_best = np.where(ok & (tmin < t_best), tmin, t_best)
    z = t_best  # d_cam has unit z, so the ray parameter is the depth
    z[~np.isfinite(z) | (z > max_depth)] = 0.0
    return z


def add_depth_noise(z, rng, noise_k: float = 0.0015):
    """z += N(0, noise_k z^2) on the hit pixels (z > 0), in place, in pixel order."""
    hit = z > 0
    z[hit] += rng.normal(0.0, 1.0, hit.sum()) * noise_k * z[hit] ** 2


def backproject(depth, fx=525.0, fy=525.0, cx=319.5, cy=239.5):
    """Camera-frame points (n, 3) float32 of the valid pixels."""
    h, w = depth.shape
    v, u = np.nonzero(depth > 0)
    z = depth[v, u].astype(np.float64)
    return np.stack([(u - cx) / fx * z, (v - cy) / fy * z, z], axis=1).astype(np.float32)


def config5_frames(n_frames: int = 64, seed: int = 0, width: int = 640, height: int = 480):
    """Streaming workload: depth frames of the config-2 room along camera_path,
    depth noise N(0, 0.0015 z^2).  Returns (depth images, camera poses,
    intrinsics scale) with intrinsics scaled to the resolution."""
    from concurrent.futures import ThreadPoolExecutor
    boxes = room_boxes(seed)
    poses = camera_path(n_frames, seed)
    s = width / 640.0
    rng = rng_for(5, 1000 + seed)

    def clean(T):
        # noise-free float64 depth (the noise is drawn below, frame by frame,
        # from one generator: the same draws as rendering the frames in order)
        return render_depth_f64(T, boxes, width, height, 525.0 * s, 525.0 * s, (width - 1) / 2.0,
                                (height - 1) / 2.0)

    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        zs = list(ex.map(clean, poses))
    depths = []
    for z in zs:
        add_depth_noise(z, rng)
        depths.append(z.reshape(height, width).astype(np.float32))
    return np.stack(depths), poses, s
