"""Shared test helpers (comparison metrics against the oracle)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return os.path.join(GOLDEN, name)


def rot_err(Ta, Tb):
    """Frobenius norm of the rotation difference."""
    return float(np.linalg.norm(np.asarray(Ta)[:3, :3] - np.asarray(Tb)[:3, :3]))


def trans_err(Ta, Tb):
    return float(np.linalg.norm(np.asarray(Ta)[:3, 3] - np.asarray(Tb)[:3, 3]))


def affine_fit(a, b):
    """Least-squares affine map b ~ A a + t (fp64); returns 4x4 and rms residual."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    X = np.concatenate([a, np.ones((len(a), 1))], 1)
    M, *_ = np.linalg.lstsq(X, b, rcond=None)
    T = np.eye(4)
    T[:3, :3] = M[:3].T
    T[:3, 3] = M[3]
    res = X @ M - b
    return T, float(np.sqrt(np.mean(np.sum(res * res, 1))))
