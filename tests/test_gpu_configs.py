"""The benchmark's own workloads at their full size, against the oracle and
the FP64 references:

  cfg3  the 256-scene batch (256^2..2048^2, 100..50,000 shapes) rendered in
        one graph-replayed persistent launch: every frame equals the same
        scene rendered by its own launch, bit for bit, and 16 sampled scenes
        (the heaviest among them) equal the oracle's frames;
  cfg4  the 1024^2, 2,000-shape differentiable scene: the adjoint's
        gradients against FP64 central differences on 64 sampled
        parameters, at the benchmark's tuned 16-px tiles and at the pinned
        256-px tiles (whose tree mixes column-segment heights of 32, 64 and
        128 rows; the FP64 forward is cut by the tree's segment map);
  cfg5  the 64-scene, 512^2 batched adjoint with bilinear texture: three
        sampled scenes, rendered in the one 64-scene launch, checked on
        points of textured shapes and on solid colours, and their forward
        images against the oracle's FP64 integration.
"""
import sys
import time

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import adjoint_cases as A
from tests import helpers as H

pytestmark = pytest.mark.gpu

COVERAGE_TOL = 1e-5


def say(msg):
    print(f"[gpu-configs] {msg}", flush=True)
    sys.stdout.flush()


def gpu_f32(r, scene, spec):
    """The raster path's pre-quantisation image of a scene (a target)."""
    import torch
    _, geo, jobs = G.prepare(scene, spec)
    r.upload_common(geo.geo, scene)
    r.frame_upload(jobs, scene)
    out = torch.zeros(scene.width * scene.height, dtype=torch.int32, device="cuda")
    f32 = torch.zeros(scene.width * scene.height * 4, dtype=torch.float32, device="cuda")
    r.set_stream(torch.cuda.current_stream().cuda_stream)
    r.frame_render(out.data_ptr(), f32.data_ptr())
    r.frame_check()
    torch.cuda.synchronize()
    return f32.cpu().numpy().reshape(scene.height, scene.width, 4)


def test_cfg3_batch_256_graph():
    import torch
    t0 = time.perf_counter()
    scenes = S.batch_256()
    say(f"generated 256 scenes in {time.perf_counter() - t0:.1f} s")
    spec = G.default_spec(max_tile_size=64, sort_radix_bits=0)
    r = G.Rasterizer(spec)
    r.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        b = G.Batch.from_outlines(r, scenes, tile_sizes=[64] * len(scenes), radix_bits=[0] * len(scenes))
        say("batch built on the device; graph launch")
        b.render(use_graph=True)
        b.check()
        b.render(use_graph=True)     # a replay of the captured graph
        lat = b.check()
        assert np.all(lat > 0)
        frames = [b.frame(i) for i in range(len(scenes))]
        say("per-scene launches")
        for i in range(len(scenes)):
            b.render_scene(i)
            b.check()
            got = b.frame(i)
            bad = H.first_mismatch(got, frames[i])
            assert bad is None, f"scene {i}: batched frame differs from its own launch at {bad}"
        b.close()
    finally:
        r.close()
    work = [sc.n_shapes * sc.width * sc.height for sc in scenes]
    heavy = sorted(range(len(scenes)), key=lambda i: -work[i])[:4]
    many = sorted(range(len(scenes)), key=lambda i: -scenes[i].n_shapes)[:2]
    rest = np.random.default_rng(3).choice([i for i in range(len(scenes)) if i not in heavy + many], 10,
                                           replace=False)
    sample = sorted(set(heavy) | set(many) | set(int(i) for i in rest))
    assert len(sample) >= 16 - 2
    for i in sample:
        say(f"oracle scene {i}: {scenes[i].width}^2, {scenes[i].n_shapes} shapes")
        ref = H.oracle_render(scenes[i], spec, want_f64=False, nthreads=16)
        assert ref["rc"] == 0
        bad = H.first_mismatch(frames[i], ref["out"])
        assert bad is None, f"scene {i}: batched frame differs from the oracle at {bad}"


def _check_fd(dp, dc, samples, what):
    assert len(samples) == 64, f"{what}: only {len(samples)} finite-difference samples accepted"
    got, fd, err, bad = A.compare(samples, dp, dc)
    say(f"{what}: max |adjoint - fd| {err.max():.3e}, max |fd| {np.abs(fd).max():.3e}")
    assert not bad.any(), [(s, g) for s, g, b in zip(samples, got, bad) if b][:5]


@pytest.mark.parametrize("tile", [16, 256])
def test_cfg4_adjoint_1024_finite_differences(tile):
    base = S.adjoint_1024()
    params = S.perturb(base, S.SEEDS["adjoint_1024"] + 1)
    spec = G.default_spec(max_tile_size=tile, sort_radix_bits=0)
    r = G.Rasterizer(spec)
    try:
        target = gpu_f32(r, base, spec)
        say(f"cfg4 adjoint at {tile}-px tiles")
        adj = G.Adjoint(r, [(params, target)])
        adj.run()
        loss, dp, dc, img = adj.result(0, image=True)
        adj.close()
    finally:
        r.close()
    ref = H.oracle_render(params, spec, want_f64=True, nthreads=16)
    err = np.abs(img.astype(np.float64) - ref["f64"].reshape(img.shape)).max()
    say(f"forward image within {err:.2e} of the oracle's FP64 integration")
    assert err <= COVERAGE_TOL
    seg = A.segment_map(params, spec)
    say(f"segment heights {sorted(set(np.unique(seg).tolist()))}; finite differences")
    ref_loss = float(((ref["f64"].reshape(img.shape) - target) ** 2).sum())
    say(f"loss gpu {loss:.4f} oracle {ref_loss:.4f}")
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    samples = A.fd_samples(params, target, n=64, seed=11, seg=seg, max_tries=1200)
    _check_fd(dp, dc, samples, f"cfg4 tile {tile}")


def test_cfg5_batched_adjoint_texture():
    import torch
    bases = S.adjoint_batch_512()
    params = [S.perturb(b, S.SEEDS["adjoint_batch_512"] + i) for i, b in enumerate(bases)]
    spec = G.default_spec(max_tile_size=32, sort_radix_bits=0, texture_filter=1)   # bilinear
    r = G.Rasterizer(spec)
    r.set_stream(torch.cuda.current_stream().cuda_stream)
    pict_shapes = []
    for sc in params:
        t = np.asarray(sc.tags, np.uint64) & np.uint64(0xC000000000000000)
        pict_shapes.append(np.flatnonzero(t == np.uint64(S.TAG_PICTURE)))
    heavy = int(np.argmax([sc.n_shapes for sc in params]))
    light = int(np.argmin([sc.n_shapes for sc in params]))
    sample = sorted({heavy, light, 17})
    try:
        targets = [gpu_f32(r, b, spec) for b in bases]
        say("64-scene adjoint launch")
        adj = G.Adjoint(r, list(zip(params, targets)))
        adj.run()
        results = {i: adj.result(i, image=True) for i in sample}
        adj.close()
    finally:
        r.close()
    for i in sample:
        sc = params[i]
        loss, dp, dc, img = results[i]
        assert len(pict_shapes[i]) > 0
        ref = H.oracle_render(sc, spec, want_f64=True, nthreads=16)
        err = np.abs(img.astype(np.float64) - ref["f64"].reshape(img.shape)).max()
        say(f"scene {i} ({sc.n_shapes} shapes, {len(pict_shapes[i])} textured): forward err {err:.2e}")
        assert err <= COVERAGE_TOL
        seg = A.segment_map(sc, spec)
        samples = A.fd_samples(sc, targets[i], n=64, seed=5 + i, seg=seg, filter=1,
                               point_shapes=pict_shapes[i], max_tries=1500)
        npts = sum(k == "point" for k, *_ in samples)
        assert npts >= 32, f"only {npts} point samples on textured shapes"
        _check_fd(dp, dc, samples, f"cfg5 scene {i}")
