"""gudni_frame_render_static (the entry point that allocates nothing: its
kernels take rings, shape scratch, control words and the scene descriptor
from the library's static device array) against gudni_frame_render, through
the C ABI: the same BGRA words and pre-quantisation colours on two calls in a
row, the second with a smaller frame, plus frames that need the retry pass
and nonzero fill; sizes beyond the array's limits are errors."""
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import helpers as H
from tests.test_gpu_parity import comb_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-static] {msg}", flush=True)
    sys.stdout.flush()


def render_both(r, scene, spec=None):
    """(frame_render words, colours), (frame_render_static words, colours)."""
    import torch
    spec = spec or r.spec
    _, geo, jobs = G.prepare(scene, spec)
    r.upload_common(geo.geo, scene)
    r.frame_upload(jobs, scene)
    r.set_stream(torch.cuda.current_stream().cuda_stream)
    n = scene.width * scene.height
    out = []
    for static in (False, True):
        u = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        f = torch.full((n * 4,), -1.0, dtype=torch.float32, device="cuda")
        if static:
            r.frame_render_static(u.data_ptr(), f.data_ptr())
        else:
            r.frame_render(u.data_ptr(), f.data_ptr())
        r.frame_check()
        torch.cuda.synchronize()
        out.append((u.cpu().numpy().view(np.uint32).reshape(scene.height, scene.width),
                    f.cpu().numpy().reshape(scene.height, scene.width, 4)))
    return out


def test_static_render_equals_frame_render():
    r = G.Rasterizer()
    try:
        scenes = [S.blobs_scaled(768, 6000, seed=31), S.batch_scene(160, 300, 7), comb_scene(),
                  S.uniform_512(300, 256, 4)]
        sizes = [sc.width * sc.height for sc in scenes]
        assert sizes[1] < sizes[0], "the second call must be the smaller input"
        for sc in scenes:
            say(f"{sc.name} {sc.width}x{sc.height}, {sc.n_shapes} shapes")
            (a, fa), (b, fb) = render_both(r, sc)
            assert H.first_mismatch(b, a) is None, f"{sc.name}: static frame differs from frame_render"
            assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32)), f"{sc.name}: colours differ"
        ref = H.oracle_render(scenes[-1], want_f64=False)["out"]
        assert H.first_mismatch(b, ref) is None
    finally:
        r.close()


def test_static_render_limits():
    r = G.Rasterizer(G.default_spec(max_thresholds=8192))
    try:
        import torch
        sc = H.tiny_square()
        _, geo, jobs = G.prepare(sc, r.spec)
        r.upload_common(geo.geo, sc)
        r.frame_upload(jobs, sc)
        out = torch.zeros(sc.width * sc.height, dtype=torch.int32, device="cuda")
        with pytest.raises(G.GudniError) as ei:
            r.frame_render_static(out.data_ptr())
        assert ei.value.code == -1 and "max_thresholds" in str(ei.value)
    finally:
        r.close()
    # an empty 4096^2 canvas cut into 8-pixel tiles: 262,144 work units
    spec = G.default_spec(max_tile_size=8)
    r = G.Rasterizer(spec)
    try:
        import torch
        sc = S.build_scene(4096, 4096, [], np.zeros((0, 4), np.float32), name="empty")
        _, geo, jobs = G.prepare(sc, spec)
        r.upload_common(geo.geo, sc)
        r.frame_upload(jobs, sc)
        out = torch.zeros(sc.width * sc.height, dtype=torch.int32, device="cuda")
        with pytest.raises(G.GudniError) as ei:
            r.frame_render_static(out.data_ptr())
        assert ei.value.code == -1 and "work units" in str(ei.value)
        r.frame_render(out.data_ptr())    # the allocating entry point still renders it
        r.frame_check()
    finally:
        r.close()
