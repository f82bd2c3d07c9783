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
