"""Checks of the coarse proposer's CPU restatement (oracle/coarse.py) that
need no GPU: known answers that fix its voxel indexing and its rotation
direction (the TDF and FPS it builds on are pinned in test_oracle.py)."""
import numpy as np

from oracle import coarse as oc
from realtime_robot_amd import synth


def _target(seed=0):
    _, tgts, _ = synth.config4_batch(1, seed_offset=seed, ns_range=(3000, 4000), nt_range=(4000, 6000))
    return tgts[0]


def test_identical_clouds_score_zero_at_zero_yaw():
    """A keypoint's own occupancy lands exactly on the occupied voxels of its
    own TDF (voxel v of the +-10 cm grid is voxel v + 5 of the +-15 cm grid),
    so the score is exactly 0 at yaw 0 and no yaw does better."""
    t = _target()
    sc, ks, kt = oc.pair_scores(t, t)
    assert np.array_equal(ks, kt)
    for i in range(oc.KEYS):
        assert sc[i, i, 0] == 0.0
        assert np.argmin(sc[i, i]) == 0


def test_quarter_turn_is_found_in_the_positive_sense():
    """Source = target turned by -90 degrees about z: every keypoint pairs
    with its own image and the best yaw is step 9 (+90 degrees, source ->
    target), not step 27 (a sign error) nor 0."""
    t = _target(1)
    R = np.array([[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])  # Rz(-90)
    s = (t.astype(np.float64) @ R.T).astype(np.float32)
    sc, ks, kt = oc.pair_scores(s, t)
    assert np.abs(ks - (kt.astype(np.float64) @ R.T)).max() == 0.0  # FPS commutes with the rotation
    best = np.array([np.argmin(sc[i, i]) for i in range(oc.KEYS)])
    assert np.mean(best == 9) >= 0.9, best
    assert np.median([sc[i, i, 9] for i in range(oc.KEYS)]) < 0.1
    assert np.median([sc[i, i, 27] for i in range(oc.KEYS)]) > 1.0


def test_occupancy_count_filter_rejects_unbalanced_pairs():
    grid = np.zeros((oc.TDFD,) * 3, np.float32)
    vox = np.arange(100)
    assert np.all(oc.score(vox, grid, 40) == oc.REJECT)      # 100 > 2 x 40
    assert np.all(oc.score(vox, grid, 210) == oc.REJECT)     # 210 > 2 x 100
    assert np.all(oc.score(vox, grid, 60) < oc.REJECT)
