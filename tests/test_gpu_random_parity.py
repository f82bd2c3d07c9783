"""A seeded sweep of small scenes and specs: blob, glyph-outline and uniform
scenes of random size and shape count, random substance mixes (solid,
gradient, picture), tile size, sort radix, texture filter and coverage mode.
Every frame's BGRA words must equal the oracle's render under the same spec,
and pre-quantisation colours must be within 1e-5 of its FP64 integration."""
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import helpers as H
from tests.test_gpu_parity import check_frames

pytestmark = pytest.mark.gpu

N_CASES = 64
REFERENCE_MODE_TOL = 5e-5


def say(msg):
    print(f"[gpu-random] {msg}", flush=True)
    sys.stdout.flush()


def case(i):
    rng = np.random.default_rng(9000 + i)
    kind = i % 3
    size = int(rng.choice([48, 96, 160, 256, 300]))
    if kind == 0:
        scene = S.batch_scene(size, int(rng.integers(5, 400)), 100 + i, gradient_share=int(rng.integers(0, 4)),
                              picture_share=int(rng.integers(0, 3)))
    elif kind == 1:
        scene = S.blobs_scaled(size, int(rng.integers(20, 800)), 200 + i, glyph_share=float(rng.uniform(0.05, 0.5)))
    else:
        scene = S.uniform_512(int(rng.integers(5, 120)), size, 300 + i)
    spec = G.default_spec(max_tile_size=int(rng.choice([16, 32, 64, 256])),
                          sort_radix_bits=int(rng.integers(0, 6)),
                          texture_filter=int(rng.integers(0, 2)),
                          coverage_mode=int(rng.integers(0, 2)))
    return scene, spec


@pytest.mark.parametrize("i", range(N_CASES))
def test_random_scene_matches_oracle(i):
    scene, spec = case(i)
    say(f"case {i}: {scene.name} {scene.width}x{scene.height} {scene.n_shapes} shapes, tile "
        f"{spec.max_tile_size} radix {spec.sort_radix_bits} filter {spec.texture_filter} "
        f"mode {spec.coverage_mode}")
    ref = H.oracle_render(scene, spec)
    assert ref["rc"] == 0
    r = G.Rasterizer(spec)
    try:
        u32, f32 = H.gpu_render_level2(r, scene, spec)
    finally:
        r.close()
    if spec.coverage_mode == 0:
        check_frames(scene, u32, f32, ref)
    else:
        # the reference's own mid-point arithmetic: bit for bit with the
        # oracle's FP32 colours; its distance to the FP64 integration can pass
        # 1e-5 in tall segments (the reference rounds y0 + (y1 - y0)/2)
        assert H.first_mismatch(u32, ref["out"]) is None
        np.testing.assert_array_equal(f32.view(np.uint32), np.asarray(ref["f32"], np.float32).reshape(f32.shape).view(np.uint32))
        assert np.abs(f32.astype(np.float64) - ref["f64"]).max() <= REFERENCE_MODE_TOL


N_CUT_CASES = 12
SENTINEL = np.uint32(0xDEADBEEF)     # alpha 0xDE: never a pixel the path writes


@pytest.mark.parametrize("i", range(N_CUT_CASES))
def test_random_level1_jobs_with_cut_tile_boxes(i):
    """Level 1 with hand-made TileInfo: the sweep's scenes and specs, every
    job's tile boxes cut at random to 1..height rows (half of the tiles).  A
    column is active while its origin lies in the cut box and writes its full
    segment regardless (initTileState / isActiveThread, Kernels.cl:1692-1722).
    Per job, the pixels gudni_raster_job changes in a frame of sentinels are
    exactly those the oracle writes, with the oracle's BGRA words."""
    from oracle import oracle as O
    scene, spec = case(i)
    rng = np.random.default_rng(7000 + i)
    _, geo, jobs = G.prepare(scene, spec)
    ogeo = O.Geometry(scene)
    assert np.array_equal(np.frombuffer(ogeo.geo.tobytes(), np.uint8), np.frombuffer(geo.geo.tobytes(), np.uint8))
    ospec = H.oracle_spec(spec)
    say(f"cut boxes {i}: {scene.name} {scene.width}x{scene.height}, {jobs.n_jobs} jobs, tile {spec.max_tile_size}")
    r = G.Rasterizer(spec)
    try:
        r.upload_common(geo.geo, scene)
        for j in range(jobs.n_jobs):
            tiles, shapes = jobs.job(j)
            t = np.frombuffer(np.ascontiguousarray(tiles).tobytes(), H.TILE_DT).copy()
            bh = t["box"][:, 3] - t["box"][:, 1]
            cut = rng.random(t.size) < 0.5
            t["box"][cut, 3] = t["box"][cut, 1] + rng.integers(1, bh[cut] + 1)
            cut_tiles = t.view(np.uint8)
            shapes = np.ascontiguousarray(shapes, np.uint8)

            class OneJob:
                n_jobs = jobs.n_jobs

                def job(self, _):
                    return cut_tiles, shapes

            rc, ref, _ = O.raster_job_dump(scene, ospec, ogeo, OneJob(), j, cap=1)
            assert rc == 0
            out = np.full(scene.width * scene.height, SENTINEL, np.uint32)
            r.raster_job(shapes, cut_tiles, scene.background, scene.width, scene.height, 0,
                         jobs.n_jobs - 1 - j, out)
            frame = out.reshape(scene.height, scene.width)
            wrote = frame != SENTINEL
            np.testing.assert_array_equal(wrote, ref != 0, err_msg=f"job {j}: written pixels")
            assert np.array_equal(frame[wrote], ref[wrote]), f"job {j}: {H.first_mismatch(np.where(wrote, frame, 0), ref)}"
    finally:
        r.close()
