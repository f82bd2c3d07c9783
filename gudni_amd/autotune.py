"""Autotuner for tile size and sort radix.

The reference's tile size is whatever the device reports as its maximum
work-group size (determineRasterSpec, OpenCL/Setup.hs:71-87); any power of two
up to 256 gives a well-defined tile tree, and every column still runs the
same threshold / sort / coverage semantics.  Smaller tiles split crowded
regions further (shorter columns, fewer shapes per tile, more waves) at the
cost of more shape references.  The sort radix is the digit width of the
radix pass ahead of each column's insertion sort: it pays where columns hold
many records, and every width yields the same permutation.

`tune_batch` picks one tile size and radix for a whole batch by timing the
whole batch (throughput).  `tune` groups the scenes of a batch by canvas
side and shape count (latency: each group's own best), renders a
few scenes of each group, repeated until the batch holds `fill_pixels`
pixels (so the device is as full as in the real batch and the timing
measures throughput, not the latency of a few scenes), with each candidate
tile size (at the default radix), then with each candidate radix at the
chosen tile size, and keeps the fastest per group.  The timed batches are
built on the device from the outlines.  `tune_frame` does the same for one
frame with a caller-supplied timing function.
"""
from __future__ import annotations

import json
import time

import numpy as np

from . import Batch, default_spec, prepare

CANDIDATES = (32, 64, 128, 256)
# the adjoint's scenes are sparse: a 16-pixel tile still has work for every lane
ADJOINT_CANDIDATES = (16, 32, 64, 128, 256)
RADIX_CANDIDATES = (0, 3, 5)


def bucket(scene) -> str:
    side = max(scene.width, scene.height)
    sb = "s512" if side <= 512 else ("s1024" if side <= 1024 else "s2048")
    n = scene.n_shapes
    nb = "n1k" if n < 1000 else ("n10k" if n < 10_000 else "n50k")
    return f"{sb}-{nb}"


def _time_render(b, reps: int) -> float:
    import torch
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


def search(time_fn, tiles=CANDIDATES, radices=RADIX_CANDIDATES, default_radix: int | None = None):
    """Coordinate search: the fastest tile size at the default radix, then the
    fastest radix at that tile size.  time_fn(tile, radix) -> ms.  Returns
    (tile, radix, {"tile,radix": ms})."""
    r0 = default_spec().sort_radix_bits if default_radix is None else default_radix
    timings = {}
    for ts in tiles:
        timings[(ts, r0)] = time_fn(ts, r0)
    best_tile = min(tiles, key=lambda ts: timings[(ts, r0)])
    for rb in radices:
        if (best_tile, rb) not in timings:
            timings[(best_tile, rb)] = time_fn(best_tile, rb)
    best = min(((best_tile, rb) for rb in set(radices) | {r0}), key=lambda k: timings[k])
    return best[0], best[1], {f"{t},{r}": round(ms, 3) for (t, r), ms in timings.items()}


def tune(rasterizer, scenes, per_bucket: int = 3, reps: int = 3, candidates=CANDIDATES,
         radices=RADIX_CANDIDATES, log=None, fill_pixels: int = 16 << 20) -> dict:
    """Returns {bucket: {"tile": size, "radix": bits}} measured on the
    rasterizer's device."""
    groups: dict[str, list] = {}
    for sc in scenes:
        groups.setdefault(bucket(sc), []).append(sc)
    table = {}
    for key, members in sorted(groups.items()):
        sample = members[:per_bucket]
        pixels = sum(sc.width * sc.height for sc in sample)
        filled = sample * max(1, -(-fill_pixels // pixels))

        def time_fn(ts, rb):
            n = len(filled)
            return _time_render(Batch.from_outlines(rasterizer, filled, [ts] * n, [rb] * n), reps)

        ts, rb, timings = search(time_fn, candidates, radices)
        table[key] = {"tile": ts, "radix": rb}
        if log:
            log(f"autotune {key}: {json.dumps(timings)} -> tile {ts}, radix {rb}")
    return table


def tune_batch(rasterizer, scenes, reps: int = 2, tiles=CANDIDATES, radices=RADIX_CANDIDATES, log=None):
    """One tile size and radix for every scene of a batch, timed on the whole
    batch (built on the device): the throughput choice.  A small tile
    shortens each scene's latency but re-reads more geometry, and in a
    full batch the device is busy anyway.  Returns (tile, radix,
    {"tile,radix": ms})."""
    n = len(scenes)

    def time_fn(ts, rb):
        return _time_render(Batch.from_outlines(rasterizer, scenes, [ts] * n, [rb] * n), reps)

    ts, rb, timings = search(time_fn, tiles, radices)
    if log:
        log(f"autotune whole batch: {json.dumps(timings)} -> tile {ts}, radix {rb}")
    return ts, rb, timings


def tune_frame(time_fn, tiles=CANDIDATES, radices=(0, 3, 4, 5), default_radix: int = 0):
    """Tile size (timed at default_radix) and then radix for one frame.
    time_fn(tile, radix) -> ms.  Returns (tile, radix, {"tile,radix": ms})."""
    return search(time_fn, tiles, radices, default_radix)


def prepare_batch(scenes, table: dict | None = None, default_tile: int = 256, default_radix: int | None = None):
    """(scene, Geometry, RasterJobs) per scene with the tuned tile size and
    radix (per group from `table`, else the defaults); returns (items, tile
    sizes, radices, host seconds)."""
    items, tiles, radices = [], [], []
    t0 = time.perf_counter()
    r0 = default_spec().sort_radix_bits if default_radix is None else default_radix
    for sc in scenes:
        choice = (table or {}).get(bucket(sc), {"tile": default_tile, "radix": r0})
        spec = default_spec(max_tile_size=choice["tile"], sort_radix_bits=choice["radix"])
        items.append((sc,) + prepare(sc, spec)[1:])
        tiles.append(choice["tile"])
        radices.append(choice["radix"])
    return items, np.array(tiles), np.array(radices), time.perf_counter() - t0
