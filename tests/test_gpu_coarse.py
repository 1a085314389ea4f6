"""The keypoint + TDF + yaw-search coarse proposer (rtr_coarse_*) against its
CPU restatement (oracle/coarse.py): keypoints and occupied voxels bit-exact,
keypoint-pair scores to float summation order, and the proposals it makes on
gravity-aligned pairs."""
import numpy as np
import pytest

from oracle import coarse as oc
from realtime_robot_amd import synth
from realtime_robot_amd.coarse import CoarseProposer

pytestmark = pytest.mark.gpu


def test_keypoints_occupancy_and_scores_match_oracle():
    srcs, tgts, _ = synth.config4_yaw_batch(2, seed_offset=300, ns_range=(3000, 9000), nt_range=(5000, 12000))
    cp = CoarseProposer()
    cp.set_clouds(srcs, tgts)
    sc, yaw, kp = cp.pair_scores()
    occ = cp.occupancy()
    for p in range(2):
        ref, ks, kt = oc.pair_scores(srcs[p], tgts[p])
        assert np.array_equal(kp[2 * p], ks) and np.array_equal(kp[2 * p + 1], kt)
        for j in range(oc.KEYS):
            assert np.array_equal(occ[p][j], oc.occupancy(srcs[p], ks[j]))
        best = ref.min(axis=2)
        rejected = best >= 1e38
        assert np.array_equal(sc[p] >= 1e38, rejected)
        ok = ~rejected
        rel = np.abs(sc[p][ok] - best[ok]) / np.maximum(best[ok], 1e-3)
        print(p, "pairs kept", ok.sum(), "max rel score error", rel.max())
        assert rel.max() < 1e-5
        # the best yaw agrees wherever the oracle's best is not a near-tie
        srt = np.sort(ref, axis=2)
        clear = ok & (srt[:, :, 1] - srt[:, :, 0] > 1e-3 * np.maximum(srt[:, :, 0], 1e-3))
        assert np.array_equal(yaw[p][clear], np.argmin(ref, axis=2)[clear])


def test_coarse_proposals_on_gravity_aligned_pairs():
    srcs, tgts, gts = synth.config4_yaw_batch(16, seed_offset=310, ns_range=(3000, 20000), nt_range=(10000, 60000))
    cp = CoarseProposer()
    cp.set_clouds(srcs, tgts)
    T, s = cp.propose()
    rot = [np.degrees(np.arccos(np.clip((np.trace(T[i][:3, :3].T @ gts[i][:3, :3]) - 1) / 2, -1, 1)))
           for i in range(16)]
    rmse = [synth.rmse_between(T[i], gts[i], srcs[i]) for i in range(16)]
    print("coarse: rotation error deg", np.round(rot, 1), "rmse m", np.round(rmse, 3), f"{cp.last_ms():.2f} ms")
    assert np.all(np.abs(np.linalg.det(T[:, :3, :3]) - 1) < 1e-9)
    assert np.mean(np.array(rot) < 15.0) >= 0.75


def test_ablation_learned_vs_reference_method():
    """The learned proposer against the reference's own coarse method and
    against no proposer, each followed by the same ICP refine: on full-SO(3)
    pairs the learned path registers more pairs than either; on gravity-
    aligned pairs the yaw search beats starting without a proposer."""
    from realtime_robot_amd import ablation
    out = ablation.run(pairs=24, seed_offset=500, verbose=False)
    for wl in ("so3", "yaw"):
        print(wl, {k: (v["within_1mm"], round(v["proposal_rot_err_deg_median"], 1)) for k, v in out[wl].items()})
    so3, yaw = out["so3"], out["yaw"]
    assert so3["learned"]["within_1mm"] > so3["coarse"]["within_1mm"]
    assert so3["learned"]["within_1mm"] > so3["none"]["within_1mm"]
    assert yaw["coarse"]["within_1mm"] > yaw["none"]["within_1mm"]
