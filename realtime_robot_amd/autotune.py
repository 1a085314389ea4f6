"""Autotuner for the batched registration engine.

Picks the grid cell size (as a fraction of the max correspondence distance),
the cell budget per target point, the work-item size, the LDS candidate
tiles on/off and the persistent workgroups per CU by timing the real batch
(HIP events around the persistent ICP kernel and the grid build, on the
engine's own stream).  Results are cached in a JSON file keyed by the GPU
architecture, the correspondence metric and coarse size buckets of the
batch, so later runs with similar batches start from the tuned setting.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import _lib as L

CACHE = os.environ.get("RTR_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "realtime_robot_amd", "tune.json"))

# (cell_factor, cells_per_point, item_pts, blocks_per_cu, use_tile[, retire])
# blocks_per_cu also picks the search kernel: 1-2 the 8-deep one (2 waves per
# SIMD), 4 the one built for 4 waves per SIMD, else 3 waves per SIMD
# retire: workgroups beyond one per CU leave once fewer than this fraction of
# the problems are unfinished (0: never; rtr_batch_set_retire)
DEFAULT = (0.25, 2.0, 1024, 0, 0)
CANDIDATES = [
    (0.25, 2.0, 1024, 0, 0),
    (0.25, 2.0, 1024, 0, 1),
    (0.35, 2.0, 1024, 0, 0),
    (0.2, 3.0, 1024, 0, 0),
    (0.25, 2.0, 512, 0, 0),
    (0.25, 2.0, 1024, 2, 0),
    (0.35, 2.0, 1024, 0, 0, 0.25),
    (0.35, 2.0, 1024, 0, 0, 0.5),
    # 4 workgroups per CU: the search kernel compiled for 4 waves per SIMD
    (0.35, 2.0, 1024, 4, 0),
    (0.25, 2.0, 1024, 4, 0),
    (0.35, 2.0, 1024, 4, 0, 0.5),
    # LDS candidate tiles (each wave stages its queries' 3x3x3 blocks)
    (0.35, 2.0, 1024, 0, 1),
    (0.25, 2.0, 512, 0, 1),
    (0.35, 2.0, 512, 2, 1),
]
# Few problems per launch (one refine per pair, a small fallback set): the
# launch is bound by each problem's per-iteration latency, so smaller work
# items (more workgroups per problem per iteration) are tried as well.
# Fewer persistent workgroups per CU than the occupancy limit are tried too:
# extra co-resident workgroups raise throughput but stretch each problem's
# per-iteration latency.
LATENCY_CANDIDATES = CANDIDATES + [
    (0.25, 2.0, 256, 0, 0),
    (0.35, 2.0, 256, 0, 0),
    (0.2, 3.0, 256, 0, 0),
    (0.35, 2.0, 512, 0, 0),
    (0.25, 2.0, 512, 2, 0),
    (0.25, 2.0, 256, 2, 0),
    (0.35, 2.0, 512, 2, 0),
    (0.2, 3.0, 512, 2, 0),
    (0.35, 2.0, 512, 1, 0),
    (0.35, 2.0, 256, 1, 0),
    (0.35, 2.0, 512, 0, 0, 0.25),
    (0.35, 2.0, 512, 0, 0, 0.5),
    (0.35, 2.0, 512, 0, 0, 0.75),
    (0.35, 2.0, 512, 2, 0, 0.5),
    (0.35, 2.0, 256, 2, 0, 0.5),
    (0.25, 2.0, 512, 0, 0, 0.5),
    (0.25, 2.0, 512, 0, 0, 0.4),
    (0.25, 2.0, 512, 0, 0, 0.6),
    (0.25, 2.0, 256, 0, 0, 0.5),
    (0.25, 2.0, 1024, 0, 0, 0.5),
    (0.25, 2.0, 1024, 4, 0, 0.5),
    (0.25, 2.0, 512, 4, 0, 0.5),
    (0.25, 2.0, 512, 0, 1, 0.5),
    (0.25, 2.0, 256, 2, 1, 0.5),
    # 256-point items for the late iterations only: early ones keep large
    # items (fewer partial sums and claims), the long tail spreads wider
    (0.25, 2.0, 1024, 0, 0, 0.5, 6),
    (0.25, 2.0, 1024, 0, 0, 0.5, 10),
    (0.25, 2.0, 1024, 0, 0, 0.5, 16),
    (0.25, 2.0, 512, 0, 0, 0.5, 10),
    (0.25, 2.0, 1024, 4, 0, 0.5, 10),
    (0.25, 2.0, 1024, 0, 0, 0.4, 16),
    (0.25, 2.0, 1024, 0, 0, 0.5, 13),
    (0.25, 2.0, 1024, 0, 0, 0.4, 13),
]


def _arch() -> str:
    try:
        import subprocess
        out = subprocess.run(["rocm_agent_enumerator"], capture_output=True, text=True, timeout=10).stdout.split()
        gfx = [a for a in out if a.startswith("gfx") and a != "gfx000"]
        return gfx[0] if gfx else "unknown"
    except Exception:  # noqa: BLE001 - the key only needs to be stable
        return "unknown"


def bucket_key(n_problems: int, mean_src: float, mean_tgt: float, mode: int) -> str:
    b = lambda v: int(round(math.log2(max(v, 1.0))))  # noqa: E731
    return f"{_arch()}|mode{mode}|p{b(n_problems)}|s{b(mean_src)}|t{b(mean_tgt)}"


def load_cache() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache: dict) -> None:
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def apply(batch, cfg) -> None:
    cf, cpp, item, bpc, tile = cfg[:5]
    batch.set_tuning(cf, cpp, item, bpc, tile, cfg[5] if len(cfg) > 5 else 0.0, int(cfg[6]) if len(cfg) > 6 else 0)


def tune(batch, guesses=None, candidates=None, reps: int = 2, key: str | None = None, use_cache: bool = True,
         verbose: bool = False):
    """Time every candidate on `batch` (clouds, problems and params already
    set) and leave the fastest applied.  Returns (config, ms, table)."""
    candidates = candidates or CANDIDATES
    cache = load_cache() if (use_cache and key) else {}
    if key and key in cache:
        cfg = tuple(cache[key]["config"])
        apply(batch, cfg)
        return cfg, cache[key]["ms"], []
    res = (L.Result * batch.n_problems)()
    batch.set_profiling(True)
    table = []
    try:
        for cfg in candidates:
            apply(batch, cfg)
            batch.align_raw(guesses, res)  # first run of a layout allocates/captures
            ms = []
            for _ in range(reps):
                batch.align_raw(guesses, res)
                t = batch.timing()
                ms.append(t["ms_grid"] + t["ms_icp"])
            table.append((cfg, min(ms)))
            if verbose:
                print(f"[autotune] {cfg}: {min(ms):.2f} ms", flush=True)
    finally:
        batch.set_profiling(False)
    best, best_ms = min(table, key=lambda x: x[1])
    apply(batch, best)
    if key and use_cache:
        cache[key] = {"config": list(best), "ms": best_ms, "when": time.time()}
        try:
            save_cache(cache)
        except OSError:
            pass
    return best, best_ms, table
