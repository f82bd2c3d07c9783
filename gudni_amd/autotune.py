"""Tile-size autotuner for batched rendering.

The reference's tile size is whatever the device reports as its maximum
work-group size (determineRasterSpec, OpenCL/Setup.hs:71-87); any power of two
up to 256 gives a well-defined tile tree, and every column still runs the
same threshold / sort / coverage semantics.  Smaller tiles split crowded
regions further (shorter columns, fewer shapes per tile) at the cost of more
shape references.  This tuner groups scenes by canvas side and shape count,
renders a few scenes of each group as a batch with each candidate tile size,
and keeps the fastest per group.
"""
from __future__ import annotations

import json
import time

import numpy as np

from . import Batch, default_spec, prepare

CANDIDATES = (64, 128, 256)


def bucket(scene) -> str:
    side = max(scene.width, scene.height)
    sb = "s512" if side <= 512 else ("s1024" if side <= 1024 else "s2048")
    n = scene.n_shapes
    nb = "n1k" if n < 1000 else ("n10k" if n < 10_000 else "n50k")
    return f"{sb}-{nb}"


def _time_batch(rasterizer, items, reps: int) -> float:
    import torch
    b = Batch(rasterizer, items)
    try:
        b.render()
        b.check()
        stream = torch.cuda.current_stream()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(reps):
            b.render()
        stop.record(stream)
        b.check()
        return start.elapsed_time(stop) / reps
    finally:
        b.close()


def tune(rasterizer, scenes, per_bucket: int = 3, reps: int = 3, candidates=CANDIDATES,
         log=None) -> dict:
    """Returns {bucket: tile_size} measured on the rasterizer's device."""
    groups: dict[str, list] = {}
    for sc in scenes:
        groups.setdefault(bucket(sc), []).append(sc)
    table = {}
    for key, members in sorted(groups.items()):
        sample = members[:per_bucket]
        best, best_ms = None, None
        timings = {}
        for ts in candidates:
            spec = default_spec(max_tile_size=ts)
            items = [(sc,) + prepare(sc, spec)[1:] for sc in sample]
            ms = _time_batch(rasterizer, items, reps)
            timings[ts] = ms
            if best_ms is None or ms < best_ms:
                best, best_ms = ts, ms
        table[key] = best
        if log:
            log(f"autotune {key}: {json.dumps(timings)} -> {best}")
    return table


def prepare_batch(scenes, table: dict | None = None, default_tile: int = 256):
    """(scene, Geometry, RasterJobs) per scene with the tuned tile size."""
    items, chosen = [], []
    t0 = time.perf_counter()
    for sc in scenes:
        ts = (table or {}).get(bucket(sc), default_tile)
        spec = default_spec(max_tile_size=ts)
        items.append((sc,) + prepare(sc, spec)[1:])
        chosen.append(ts)
    return items, np.array(chosen), time.perf_counter() - t0
