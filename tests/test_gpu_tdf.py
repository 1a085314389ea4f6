"""GPU truncated distance field (rtr_tdf, rtr_tdf_points) against the
brute-force oracle of kernel.cu:9-31's semantics: bit-identical."""
import numpy as np
import pytest

import oracle
import realtime_robot_amd as R
from realtime_robot_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,n_occ,trunc2", [((30, 30, 30), 600, 900.0), ((64, 48, 40), 300, 900.0),
                                               ((7, 130, 3), 20, 50.0), ((1100, 3, 2), 5, 1e6),
                                               ((16, 16, 16), 0, 900.0)])
def test_tdf_matches_brute_force(dims, n_occ, trunc2):
    rng = np.random.default_rng(sum(dims) + n_occ)
    occ = rng.integers(-2, np.array(dims) + 2, size=(n_occ, 3)).astype(np.int32)   # some outside the grid
    got = R.tdf(occ, dims, trunc2)
    ref = oracle.tdf(occ, dims, trunc2)
    assert np.array_equal(got, ref)


def test_tdf_of_key_point_box_like_the_reference():
    """The reference's use: the occupied 1 cm voxels of a cloud inside a
    0.3 m box around a key point (key_point.h:287-308), 30^3 grid, cap 900."""
    src, tgt, _ = synth.config1_pair(0)
    center = tgt[1234].astype(np.float64)
    lo = (center - 0.15).astype(np.float32)
    got = R.tdf_points(tgt, lo, 0.01, (30, 30, 30), 900.0)
    f = (tgt - lo) * (np.float32(1.0) / np.float32(0.01))
    inside = np.all((f >= 0) & (f < 30), axis=1)
    occ = np.unique(f[inside].astype(np.int32), axis=0)
    ref = oracle.tdf(occ, (30, 30, 30), 900.0)
    assert len(occ) > 50 and np.array_equal(got, ref)
