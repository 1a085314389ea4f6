"""C ABI and host-side checks that need no GPU: exported symbols, PCD I/O,
the library's own geometry code run on the host (rtr_selfcheck_*), the C++
header, and the error path when no device is present."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import realtime_robot_amd as R
from helpers import GOLDEN, golden
from pcd_np import read_pcd, read_xyz
from realtime_robot_amd import _lib as L
from realtime_robot_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu_present():
    try:
        import torch  # noqa: F401  (device count only; initialises nothing)
        return torch.cuda.device_count() > 0
    except Exception:  # noqa: BLE001
        return False


def test_every_declared_symbol_is_exported():
    hdr = open(os.path.join(ROOT, "include", "rtr.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(rtr_\w+)\(", hdr, re.M))
    assert len(declared) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (rtr_\w+)", out))
    assert declared <= exported, sorted(declared - exported)
    assert set(L.EXPORTED) <= exported


def test_default_params_are_pcl_defaults():
    p = L.default_params()
    assert p.max_iterations == 10 and p.min_correspondences == 3 and p.mode == 0
    assert p.max_correspondence_distance == np.sqrt(np.finfo(np.float64).max)
    assert p.transformation_epsilon == 0.0 and p.euclidean_fitness_epsilon == -np.finfo(np.float64).max


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(GOLDEN) if f.endswith(".pcd")))
def test_pcd_reader_matches_numpy_reader(name):
    xyz, nrm = R.load_pcd(golden(name))
    d = read_pcd(golden(name))
    assert len(xyz) == d["points"]
    assert np.array_equal(xyz, read_xyz(golden(name)))
    if "normal_x" in d["fields"]:
        ref = np.stack([d["data"]["normal_x"], d["data"]["normal_y"], d["data"]["normal_z"]], 1)
        assert np.array_equal(nrm, ref.astype(np.float32), equal_nan=True)  # the file holds NaN normals
    else:
        assert nrm is None


def test_pcd_ascii_writer_reproduces_reference_output_byte_for_byte(tmp_path):
    """mcloud.pcd was written by pcl::io::savePCDFileASCII (function.h:126)."""
    xyz, _ = R.load_pcd(golden("mcloud.pcd"))
    out = tmp_path / "m.pcd"
    R.save_pcd_ascii(str(out), xyz)
    assert out.read_bytes() == open(golden("mcloud.pcd"), "rb").read()


def test_pcd_errors():
    with pytest.raises(R.RtrError) as e:
        R.load_pcd("/nonexistent/file.pcd")
    assert e.value.code == L.ERR_IO


_PCD_HEAD = "VERSION 0.7\nFIELDS {f}\nSIZE {s}\nTYPE {t}\nCOUNT {c}\nWIDTH {n}\nHEIGHT 1\nPOINTS {n}\nDATA {d}\n"


@pytest.mark.parametrize("fields,sizes,types,counts,why", [
    ("x y z", "4 4 4", "F F F", "1 1", "COUNT shorter than FIELDS"),
    ("x y z", "4 4", "F F F", "1 1 1", "SIZE shorter than FIELDS"),
    ("x y z", "4 4 4", "F F", "1 1 1", "TYPE shorter than FIELDS"),
    ("x y z", "2 4 4", "F F F", "1 1 1", "F with SIZE 2"),
    ("x y z", "4 4 3", "F F F", "1 1 1", "SIZE 3"),
    ("x y z", "4 4 4", "F F F", "1 0 1", "COUNT 0"),
    ("x y z", "4 4 4", "F F F", "1 -2 1", "negative COUNT"),
    ("x y z", "4 4 -4", "F F F", "1 1 1", "negative SIZE"),
    ("x y z", "4 4 4", "F F Q", "1 1 1", "unknown TYPE"),
])
def test_pcd_malformed_headers_are_rejected(tmp_path, fields, sizes, types, counts, why):
    """Header fields that disagree, or sizes that do not match a PCL scalar
    type, are a format error, never a read past the payload."""
    for data in ("binary", "ascii"):
        path = tmp_path / f"bad_{data}.pcd"
        body = b"\x00" * 48 if data == "binary" else b"0 0 0\n0 0 0\n"
        path.write_bytes(_PCD_HEAD.format(f=fields, s=sizes, t=types, c=counts, n=2, d=data).encode() + body)
        with pytest.raises(R.RtrError) as e:
            R.load_pcd(str(path))
        assert e.value.code == L.ERR_FORMAT, why


def test_pcd_binary_truncation_and_huge_counts(tmp_path):
    ok = tmp_path / "ok.pcd"
    xyz = np.arange(12, dtype=np.float32).reshape(4, 3)
    ok.write_bytes(_PCD_HEAD.format(f="x y z", s="4 4 4", t="F F F", c="1 1 1", n=4, d="binary").encode()
                   + xyz.tobytes())
    got, _ = R.load_pcd(str(ok))
    assert np.array_equal(got, xyz)
    # 8-byte integer fields are a PCL scalar type and are read as such
    i8 = tmp_path / "i8.pcd"
    rec = np.zeros(2, dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("label", "<i8")])
    rec["x"], rec["label"] = [1.5, -2.0], [-(1 << 40), 7]
    i8.write_bytes(_PCD_HEAD.format(f="x y z label", s="4 4 4 8", t="F F F I", c="1 1 1 1", n=2,
                                    d="binary").encode() + rec.tobytes())
    got, _ = R.load_pcd(str(i8))
    assert np.array_equal(got[:, 0], [1.5, -2.0])
    short = tmp_path / "short.pcd"
    short.write_bytes(_PCD_HEAD.format(f="x y z", s="4 4 4", t="F F F", c="1 1 1", n=4, d="binary").encode()
                      + xyz.tobytes()[:40])
    with pytest.raises(R.RtrError) as e:
        R.load_pcd(str(short))
    assert e.value.code == L.ERR_FORMAT
    # a point count whose payload size overflows size_t is truncation, not a wrap
    huge = tmp_path / "huge.pcd"
    huge.write_bytes(_PCD_HEAD.format(f="x y z", s="8 8 8", t="F F F", c="1 1 1", n=(1 << 62) + 1,
                                      d="binary").encode() + xyz.tobytes())
    with pytest.raises(R.RtrError):
        R.load_pcd(str(huge))


@pytest.mark.parametrize("cell,rmax", [(0.05, 0.08), (0.02, None), (0.3, 0.05), (0.011, 0.2), (0.004, 0.05),
                                       (0.004, None)])
def test_host_grid_search_matches_brute_force(cell, rmax):
    rng = np.random.default_rng(1)
    t = rng.random((3000, 3)).astype(np.float32)
    q = (rng.random((1500, 3)) * 1.6 - 0.3).astype(np.float32)
    i1, d1 = R.selfcheck_nn(t, q, cell, 1e30 if rmax is None else rmax)
    i2, d2 = oracle.nn_brute(q, t, rmax)
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)


def test_host_grid_search_ties_nans_and_empty_target():
    rng = np.random.default_rng(2)
    base = np.round(rng.random((2000, 3)) * 10) / 10
    t = np.concatenate([base, base[:500]]).astype(np.float32)
    t[3] = np.nan
    q = (np.round(rng.random((1000, 3)) * 20) / 20).astype(np.float32)
    q[0] = np.nan
    q[1] = [1e6, 0, 0]
    for cell in (0.02, 0.05, 0.13):
        i1, d1 = R.selfcheck_nn(t, q, cell, 0.07)
        i2, d2 = oracle.nn_brute(q, t, 0.07)
        assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    i1, _ = R.selfcheck_nn(np.zeros((0, 3), np.float32), q, 0.05, 0.1)
    assert np.all(i1 == -1)


def test_host_grid_search_on_surfaces():
    src, tgt, T = synth.config1_pair(0)
    q = oracle.transform(T, src)
    for cell in (0.006, 0.0125, 0.03, 0.0505):
        for rmax in (0.05, 0.01):
            i1, d1 = R.selfcheck_nn(tgt, q, cell, rmax)
            i2, d2 = oracle.nn_brute(q, tgt, rmax)
            assert np.array_equal(i1, i2) and np.array_equal(d1, d2)


def test_host_solver_matches_oracle():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(4)
    for k in range(4):
        a = (rng.random((400, 3)) * 2 - 1 + 3).astype(np.float32)
        T = np.eye(4)
        T[:3, :3] = Rotation.random(random_state=k).as_matrix()
        T[:3, 3] = rng.random(3)
        b = (a.astype(np.float64) @ T[:3, :3].T + T[:3, 3]).astype(np.float32)
        T1 = R.selfcheck_solve(0, a, b, anchor=(3, 3, 3))
        T2 = oracle.umeyama(a.astype(np.float64), b.astype(np.float64))
        assert np.abs(T1 - T2).max() < 1e-6
    # point-to-plane: the same linear least-squares solution as the oracle's
    a = rng.random((800, 3)).astype(np.float32)
    n = rng.normal(size=(800, 3))
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    b = (a + rng.normal(size=a.shape) * 1e-3).astype(np.float32)
    T1 = R.selfcheck_solve(1, a, b, n, anchor=(0.5, 0.5, 0.5))
    s = a.astype(np.float64) - 0.5
    d = b.astype(np.float64) - 0.5
    A = np.concatenate([np.cross(s, n.astype(np.float64)), n], 1)
    x = np.linalg.solve(A.T @ A, A.T @ np.sum(n * (d - s), 1))
    assert np.abs(T1 - oracle.euler_to_T(x, (0.5, 0.5, 0.5))).max() < 1e-6


def test_host_rigid_solver_pinned_on_hard_cases():
    """The rigid fit (Horn's quaternion: the largest eigenpair of his 4x4 by
    its characteristic polynomial, Jacobi when that eigenvalue is nearly
    repeated) against the oracle's SVD Umeyama, moments about the centroid:
    general, planar and thin clouds, rotations from 1e-7 rad to nearly pi,
    with and without noise (bound 5e-7: the per-point moment terms are FP32,
    as in the kernels, which leaves up to 2.4e-7 at large rotations; a wrong
    eigenvector is off by orders of magnitude more); a collinear cloud (rotation about the line undetermined) must
    still give a proper rotation that maps the points."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(11)
    for k in range(60):
        kind = k % 3
        a = rng.random((300, 3)) * 2 - 1
        if kind == 1:
            a[:, 2] = 0.0
        elif kind == 2:
            a[:, 1] *= 0.01
        a = (a + 3).astype(np.float32)
        ang = [1e-7, 1e-3, 0.5, 2.0, 3.1, 3.14159][k % 6]
        ax = rng.normal(size=3)
        T = np.eye(4)
        T[:3, :3] = Rotation.from_rotvec(ax / np.linalg.norm(ax) * ang).as_matrix()
        T[:3, 3] = rng.random(3)
        b = a.astype(np.float64) @ T[:3, :3].T + T[:3, 3] + rng.normal(size=a.shape) * (1e-3 if k % 2 else 0.0)
        b = b.astype(np.float32)
        T1 = R.selfcheck_solve(0, a, b, anchor=tuple(a.astype(np.float64).mean(0)))
        T2 = oracle.umeyama(a.astype(np.float64), b.astype(np.float64))
        assert np.abs(T1 - T2).max() < 5e-7, (k, np.abs(T1 - T2).max())
    a = (np.outer(rng.random(300), [1.0, 2.0, 3.0]) + 3).astype(np.float32)
    Rm = Rotation.from_rotvec([0.2, -0.1, 0.3]).as_matrix()
    b = (a.astype(np.float64) @ Rm.T + [0.1, 0.2, 0.3]).astype(np.float32)
    T1 = R.selfcheck_solve(0, a, b, anchor=tuple(a.astype(np.float64).mean(0)))
    Rs = T1[:3, :3]
    assert abs(np.linalg.det(Rs) - 1) < 1e-12 and np.abs(Rs @ Rs.T - np.eye(3)).max() < 1e-12
    assert np.abs(a.astype(np.float64) @ Rs.T + T1[:3, 3] - b).max() < 1e-5


@pytest.mark.skipif(_gpu_present(), reason="a device is present")
def test_no_device_is_a_clean_error():
    with pytest.raises(R.RtrError) as e:
        R.IterativeClosestPoint()
    assert e.value.code == L.ERR_DEVICE
    h = C.c_void_p()
    assert L.lib.rtr_batch_create(C.byref(h), 0) == L.ERR_DEVICE
    assert L.lib.rtr_map_create(C.byref(h), 0, C.c_float(0.02), 100) == L.ERR_DEVICE


def test_invalid_arguments_are_rejected():
    assert L.lib.rtr_icp_align(None, None, None, 0, None, None) == L.ERR_INVALID_ARG
    assert L.lib.rtr_map_integrate(None, None, 0, 0, None) == L.ERR_INVALID_ARG
    with pytest.raises(ValueError):
        R.selfcheck_nn(np.zeros((5, 2), np.float32), np.zeros((3, 3), np.float32), 0.1)


def _build_cpp(tmp_path):
    exe = tmp_path / "keypoint_icp"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "keypoint_icp.cpp"), "-o", str(exe),
                    "-L", os.path.dirname(L.LIB_PATH), "-lrtr", "-Wl,-rpath," + os.path.dirname(L.LIB_PATH)],
                   check=True)
    return exe


def test_cpp_header_compiles_links_and_does_io(tmp_path):
    exe = _build_cpp(tmp_path)
    out = tmp_path / "o.pcd"
    r = subprocess.run([str(exe), golden("mcloud.pcd"), golden("5.pcd"), str(out), "--io-only"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "source 1909 target 1909" in r.stdout
    assert out.read_bytes() == open(golden("mcloud.pcd"), "rb").read()


@pytest.mark.gpu
def test_cpp_keypoint_icp_on_device(tmp_path):
    exe = _build_cpp(tmp_path)
    out = tmp_path / "t.pcd"
    r = subprocess.run([str(exe), golden("5.pcd"), golden("mcloud.pcd"), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    moved = read_xyz(str(out))
    assert np.abs(moved - read_xyz(golden("mcloud.pcd"))).max() < 1e-5


def _affine_map(seed, radii):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    A = Q @ np.diag(radii) @ Q.T
    return np.ascontiguousarray(A), rng.standard_normal(6)


@pytest.mark.parametrize("m", [6, 8])
def test_anderson_mixing_terminates_on_affine_maps(m):
    """Anderson type-II mixing with memory >= the dimension is equivalent to
    GMRES on a linear fixed-point problem (Walker & Ni 2011), so on an affine
    map of R^6 it reaches the fixed point (I - A)^-1 c after at most 6 + 1
    mixing steps, whatever the contraction rate; the plain iteration needs
    thousands of steps at spectral radius 0.99."""
    A, c = _affine_map(11, [0.99, 0.97, 0.95, 0.9, 0.3, -0.5])
    xs = np.linalg.solve(np.eye(6) - A, c)
    x, it = np.empty(6), C.c_int()
    L.check(L.lib.rtr_selfcheck_anderson(m, L.ptr(A), L.ptr(c), 1e-12, 5000, L.ptr(x), C.byref(it)))
    assert it.value <= 12
    assert np.abs(x - xs).max() < 1e-9
    # small memory still helps, less
    x1, it1 = np.empty(6), C.c_int()
    L.check(L.lib.rtr_selfcheck_anderson(2, L.ptr(A), L.ptr(c), 1e-12, 5000, L.ptr(x1), C.byref(it1)))
    assert it.value < it1.value < 2000
    assert np.abs(x1 - xs).max() < 1e-8


def test_anderson_mixing_rejects_bad_arguments():
    A, c = _affine_map(3, [0.5] * 6)
    x, it = np.empty(6), C.c_int()
    for m in (0, 9):
        assert L.lib.rtr_selfcheck_anderson(m, L.ptr(A), L.ptr(c), 1e-12, 10, L.ptr(x), C.byref(it)) != 0


def test_rotation_chart_matches_quaternion_matrix():
    """The acceleration charts rotations by the vector part v of a unit
    quaternion with positive scalar part: R(v) is the textbook quaternion
    matrix (scipy's), and the chart inverts it; angles beyond ~174 degrees
    and |v| > 1 are refused."""
    from scipy.spatial.transform import Rotation as Rot
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        if q[3] < 0:
            q = -q
        if q[3] < 0.03:
            continue
        v = np.ascontiguousarray(q[:3])
        Rm, vb = np.empty(9), np.empty(3)
        L.check(L.lib.rtr_selfcheck_so3(L.ptr(v), L.ptr(Rm), L.ptr(vb)))
        assert np.abs(Rm.reshape(3, 3) - Rot.from_quat(q).as_matrix()).max() < 1e-14
        assert np.abs(vb - v).max() < 1e-12
    Rm, vb = np.empty(9), np.empty(3)
    for bad in ([0.8, 0.7, 0.0], [0.0, 0.0, 0.9995]):
        assert L.lib.rtr_selfcheck_so3(L.ptr(np.array(bad)), L.ptr(Rm), L.ptr(vb)) != 0


def test_result_conversion_matches_per_entry():
    """The batch results' numpy conversion equals the per-entry one."""
    from realtime_robot_amd.registration import result_dict, result_dicts
    res = (R._lib.Result * 5)()
    rng = np.random.default_rng(3)
    for i in range(5):
        for k in range(16):
            res[i].T[k] = rng.normal()
        res[i].mse, res[i].fitness_rmse, res[i].n_corr = rng.random(), rng.random(), i * 7
        res[i].iterations, res[i].converged, res[i].state, res[i].coarse_iterations = i + 3, i % 2, i % 4, i
    for a, b in zip(result_dicts(res), [result_dict(r) for r in res]):
        assert a.keys() == b.keys()
        assert np.array_equal(a["T"], b["T"])
        assert all(a[k] == b[k] for k in a if k != "T")
    assert result_dicts((R._lib.Result * 0)()) == []
