"""Shared helpers for the parity tests: run a scene through the HIP path and
through the oracle with identical inputs."""
from __future__ import annotations

import numpy as np

import gudni_amd as G
from gudni_amd import scenes as S
from oracle import oracle as O


def oracle_spec(spec: G.RasterSpec) -> O.Spec:
    """The oracle's spec for a product spec (computeDepth = log2 threadsPerTile)."""
    return O.default_spec(max_thresholds=spec.max_thresholds, max_shapes=spec.max_shapes,
                          compute_depth=int(spec.threads_per_tile).bit_length() - 1,
                          threads_per_tile=spec.threads_per_tile,
                          tiles_per_call=spec.max_tiles_per_call,
                          max_tile_size=spec.max_tile_size,
                          max_strands_per_tile=spec.max_strands_per_tile,
                          random_field_size=spec.random_field_size,
                          stochastic_factor=spec.stochastic_factor,
                          coverage_mode=spec.coverage_mode,
                          texture_filter=spec.texture_filter)


def oracle_render(scene, spec=None, want_f64=True, nthreads=8):
    spec = spec or G.default_spec()
    return O.render(scene, oracle_spec(spec), want_f32=True, want_f64=want_f64, nthreads=nthreads)


def gpu_render_level2(r: G.Rasterizer, scene, spec=None):
    """Whole frame through gudni_frame_upload / gudni_frame_render with device
    buffers from torch; returns (u32 frame, f32 colours) on the host."""
    import torch
    spec = spec or r.spec
    _, geo, jobs = G.prepare(scene, spec)
    r.upload_common(geo.geo, scene)
    r.frame_upload(jobs, scene)
    out = torch.zeros(scene.height * scene.width, dtype=torch.int32, device="cuda")
    f32 = torch.zeros(scene.height * scene.width * 4, dtype=torch.float32, device="cuda")
    r.set_stream(torch.cuda.current_stream().cuda_stream)
    r.frame_render(out.data_ptr(), f32.data_ptr())
    r.frame_check()
    torch.cuda.synchronize()
    u32 = out.cpu().numpy().view(np.uint32).reshape(scene.height, scene.width)
    return u32, f32.cpu().numpy().reshape(scene.height, scene.width, 4)


def gpu_render_level1(r: G.Rasterizer, scene, spec=None):
    spec = spec or r.spec
    _, geo, jobs = G.prepare(scene, spec)
    r.upload_common(geo.geo, scene)
    out = r.queue_jobs(jobs, scene)
    return out.reshape(scene.height, scene.width)


def first_mismatch(a, b):
    bad = np.argwhere(a != b)
    return None if bad.size == 0 else (tuple(bad[0]), a[tuple(bad[0])], b[tuple(bad[0])], len(bad))


def tiny_square():
    """tinySquare (benchmarks/GudniTests.hs:360-364): a 2x2 red square at
    (0.1, 0.1) over white."""
    return S.build_scene(8, 8, [(S.solid_tag(0), [S.rectangle_pairs(0.1, 0.1, 2.1, 2.1)])],
                         [[1, 0, 0, 1]], name="tinySquare")


TILE_DT = np.dtype([("box", "<i4", 4), ("hd", "<i2"), ("vd", "<i2"), ("col", "<i4"),
                    ("start", "<i4"), ("count", "<i4")])


def unit_columns(hd, cd, q, lanes):
    """Threads (reference column indices) of work unit q of a tile of width
    2^hd: a block of 2^cw columns by 2^sw row segments, 8 x 8 where the tile
    has 8 or more segments, blocks numbered across then down."""
    sw = max(0, min(cd - hd, 3))
    cw = 6 - sw
    if cw > hd:
        cw, sw = hd, 6 - hd
    n_x = 1 << (hd - cw)
    x = ((q % n_x) << cw) | (lanes & ((1 << cw) - 1))
    seg = ((q // n_x) << sw) | (lanes >> cw)
    return (seg << hd) | x


def host_frame_layout(spec, jobs, width, height):
    """The host build laid out as a whole frame, the way gudni_frame_upload
    does it: shapeStart absolute into the concatenated references, reversed job
    index per tile, and the work list of active tile quarters."""
    tiles = jobs.tiles.view(TILE_DT).copy()
    tile_job = np.zeros(jobs.n_tiles, np.int32)
    for j in range(jobs.n_jobs):
        t0, t1 = jobs.job_tile_offset[j], jobs.job_tile_offset[j + 1]
        tiles["start"][t0:t1] += jobs.job_shape_offset[j]
        tile_job[t0:t1] = jobs.n_jobs - 1 - j
    cd = int(spec.threads_per_tile).bit_length() - 1
    work = []
    lanes = np.arange(64)
    for t, r in enumerate(tiles):
        hd, vd = int(r["hd"]), int(r["vd"])
        dd = max(0, vd - (cd - hd))
        for q in range(spec.threads_per_tile // 64):
            col = unit_columns(hd, cd, q, lanes)
            ix = ((1 << hd) - 1) & col
            iy = (col >> hd) << dd
            box = r["box"]
            if np.any((iy < box[3] - box[1]) & (ix + box[0] < width) & (iy + box[1] < height)):
                work.append((t << 2) | q)
    return tiles, tile_job, jobs.shape_refs.copy(), np.array(work, np.uint32)
