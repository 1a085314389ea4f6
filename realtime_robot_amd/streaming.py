"""Streaming reconstruction: frame-to-map point-to-plane ICP and voxel-hash fusion.

The reference registers a model against one scan and overlays the
transformed clouds (RealTimeRobot.cpp:104-114); this is the same pair of
operations run over a stream of depth frames, each frame registered against
the map fused so far and then merged into it:

  * the frame's points are downsampled on the device (rtr_voxel_downsample);
  * the target is the map's voxel centroids with k-NN PCA normals computed on
    the device (viewpoint: the predicted camera position);
  * point-to-plane ICP (rtr_icp_*, mode point-to-plane) from a constant-
    velocity prediction of the camera pose;
  * the full frame is fused into the 2 cm voxel-hash map (rtr_map_integrate).

    python -m realtime_robot_amd.streaming --frames 64      # prints one JSON line
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .registration import IterativeClosestPointWithNormals, VoxelMap, voxel_downsample


class StreamingReconstruction:
    def __init__(self, voxel: float = 0.02, capacity: int = 8 << 20, icp_voxel: float = 0.02,
                 max_correspondence_distance: float = 0.03, max_iterations: int = 30, normal_k: int = 20,
                 device: int = 0):
        self.map = VoxelMap(voxel, capacity=capacity, device=device)
        self.icp = IterativeClosestPointWithNormals(device=device)
        self.icp.setMaxCorrespondenceDistance(max_correspondence_distance)
        self.icp.setMaximumIterations(max_iterations)
        self.icp.setTransformationEpsilon(1e-8)
        self.icp_voxel = icp_voxel
        self.normal_k = normal_k
        self.poses = []
        self.results = []
        self.timings = []

    def _predict(self):
        if len(self.poses) < 2:
            return self.poses[-1].copy()
        a, b = self.poses[-2], self.poses[-1]
        return b @ (np.linalg.inv(a) @ b)          # constant velocity

    def add_frame(self, pts_cam, T_init=None) -> np.ndarray:
        """Register camera-frame points against the map, fuse them, and
        return the camera-to-world pose.  The first frame is placed at
        T_init (identity by default) and starts the map."""
        t0 = time.perf_counter()
        if not self.poses:
            T = np.eye(4) if T_init is None else np.asarray(T_init, np.float64)
            self.results.append(None)
        else:
            guess = self._predict() if T_init is None else np.asarray(T_init, np.float64)
            src = voxel_downsample(pts_cam, self.icp_voxel)
            _, tgt, _ = self.map.extract(centroids=True)
            self.icp.setInputSource(src)
            self.icp.setInputTarget(tgt)
            self.icp.computeTargetNormals(self.normal_k, tuple(guess[:3, 3]))
            self.icp.align(guess)
            T = self.icp.getFinalTransformation()
            self.results.append(self.icp.getResult())
        self.map.integrate(pts_cam, T)
        self.poses.append(T)
        self.timings.append(time.perf_counter() - t0)
        return T


def trajectory_errors(est, gt):
    """(translation RMSE in m, max rotation error in degrees) of camera poses."""
    est, gt = np.asarray(est), np.asarray(gt)
    dt = est[:, :3, 3] - gt[:, :3, 3]
    c = (np.einsum("nij,nij->n", est[:, :3, :3], gt[:, :3, :3]) - 1.0) / 2.0
    return float(np.sqrt(np.mean(np.sum(dt * dt, axis=1)))), float(np.degrees(np.arccos(np.clip(c, -1, 1))).max())


def main(argv=None):
    from . import synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--width", type=int, default=640)
    ap.add_argument("--height", type=int, default=480)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--icp-voxel", type=float, default=0.02)
    ap.add_argument("--max-iter", type=int, default=30)
    ap.add_argument("--max-dist", type=float, default=0.03)
    ap.add_argument("--per-frame", action="store_true", help="add per-frame pose errors to the output")
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    depths, gt, _ = synth.config5_frames(args.frames, args.seed, args.width, args.height)
    s = args.width / 640.0
    clouds = [synth.backproject(d, 525.0 * s, 525.0 * s, (args.width - 1) / 2.0, (args.height - 1) / 2.0)
              for d in depths]
    print(f"[streaming] {args.frames} frames rendered in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    rec = StreamingReconstruction(icp_voxel=args.icp_voxel, max_iterations=args.max_iter,
                                  max_correspondence_distance=args.max_dist)
    rec.add_frame(clouds[0], gt[0])                 # the first frame fixes the map's frame
    for k in range(1, len(clouds)):
        rec.add_frame(clouds[k])
    ate, rot = trajectory_errors(rec.poses, gt)
    frame_ms = 1e3 * np.array(rec.timings[1:])
    out = {"frames": args.frames, "resolution": [args.width, args.height],
           "points_per_frame": float(np.mean([len(c) for c in clouds])),
           "frame_ms_mean": float(frame_ms.mean()), "frame_ms_p90": float(np.percentile(frame_ms, 90)),
           "frames_per_s": float(1e3 / frame_ms.mean()), "ate_rmse_m": ate, "max_rot_err_deg": rot,
           "map_voxels": rec.map.size(), "dropped": rec.map.dropped(),
           "icp_iterations_mean": float(np.mean([r["iterations"] for r in rec.results[1:]]))}
    if args.per_frame:
        P = np.asarray(rec.poses)
        out["trans_err_mm"] = [round(float(x) * 1e3, 2) for x in np.linalg.norm(P[:, :3, 3] - gt[:, :3, 3], axis=1)]
        c = (np.einsum("nij,nij->n", P[:, :3, :3], gt[:, :3, :3]) - 1.0) / 2.0
        out["rot_err_deg"] = [round(float(x), 3) for x in np.degrees(np.arccos(np.clip(c, -1, 1)))]
        out["iterations"] = [r["iterations"] for r in rec.results[1:]]
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
