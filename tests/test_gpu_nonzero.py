"""Nonzero fill on the GPU (an extension; the reference fills by crossing
parity only, Kernels.cl:298-304): frames, per-column records and shape
states equal the oracle's bit for bit, and coverage is within 1e-5 of its
FP64 integration, on the 512^2 cfg1 workload (half even-odd, half nonzero),
on scenes where the two rules differ (overlapping outlines traced the same
way round, a pentagram, random self-crossing polygons), on the device
serialiser's and the batch's paths, and under random tile sizes and sort
radices."""
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import helpers as H
from tests.test_gpu_parity import check_frames, dump_parity
from tests.test_oracle_nonzero import rect_pts, reverse_outline, star_points

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-nonzero] {msg}", flush=True)
    sys.stdout.flush()


def rule_scenes():
    """Scenes whose frames differ between the fill rules."""
    out = []
    a, b = (4, 5.25, 30, 33.5), (17, 12.75, 41, 44.0)
    same = [S.straight_pairs(rect_pts(*a)), S.straight_pairs(rect_pts(*b))]
    opp = [S.straight_pairs(rect_pts(*a)), S.straight_pairs(rect_pts(*b, reverse=True))]
    out.append(S.build_scene(48, 48, [(S.nonzero(S.solid_tag(0)), same),
                                      (S.solid_tag(1), [S.rectangle_pairs(20.5, 2.5, 46.5, 20.5)]),
                                      (S.nonzero(S.solid_tag(2)), opp)],
                             [[0, 0, 0, 1], [0.9, 0.2, 0.1, 0.5], [0.1, 0.8, 0.3, 0.6]], name="rects"))
    star = S.straight_pairs(star_points(32, 33, 27))
    out.append(S.build_scene(64, 64, [(S.nonzero(S.solid_tag(0)), [star]),
                                      (S.solid_tag(1), [reverse_outline(star)])],
                             [[0.2, 0.2, 0.9, 0.8], [0.9, 0.9, 0.1, 0.4]], name="pentagram"))
    rng = np.random.default_rng(8)
    shapes, cols = [], []
    for i in range(40):
        k = int(rng.integers(4, 12))
        pts = rng.uniform(0, 200, (k, 2)).astype(np.float32)
        ctl = pts + rng.normal(0, 15, (k, 2)).astype(np.float32)
        pairs = np.concatenate([pts, ctl], axis=1)
        tag = S.solid_tag(i)
        shapes.append((S.nonzero(tag) if i % 3 else tag, [pairs, pairs[::-1].copy()] if i % 5 == 0 else [pairs]))
        cols.append([rng.uniform(), rng.uniform(), rng.uniform(), rng.uniform(0.3, 1.0)])
    out.append(S.build_scene(200, 200, shapes, cols, name="scribbles"))
    return out


def test_cfg1_half_nonzero_frame(rasterizer):
    """The 512^2, 1000-shape workload with odd shapes nonzero: whole frame
    through Level 2 and every job through Level 1."""
    scene = S.uniform_512()
    assert sum(int(t) & S.TAG_NONZERO != 0 for t in scene.tags) == 500
    ref = H.oracle_render(scene)
    assert ref["rc"] == 0
    say("cfg1 half nonzero, level 2")
    u32, f32 = H.gpu_render_level2(rasterizer, scene)
    check_frames(scene, u32, f32, ref)
    say("cfg1 half nonzero, level 1")
    got = H.gpu_render_level1(rasterizer, scene)
    assert H.first_mismatch(got, ref["out"]) is None
    # the fill rule changes the frame: the all-even-odd copy differs
    eo = H.oracle_render(S.uniform_512(fill="evenodd"), want_f64=False)["out"]
    assert (eo != ref["out"]).mean() > 0.05


def test_cfg1_half_nonzero_records(rasterizer):
    """Per-column records (with the fill-rule header bits), queue lengths,
    shape bits, initial stacks and shape indices of two jobs."""
    dump_parity(rasterizer, S.uniform_512(), lambda n: (0, n // 2))


@pytest.mark.parametrize("mode", [0, 1])   # accurate, reference coverage
def test_rule_scenes(mode):
    spec = G.default_spec(coverage_mode=mode)
    r = G.Rasterizer(spec)
    try:
        for scene in rule_scenes():
            say(f"{scene.name} mode {mode}")
            ref = H.oracle_render(scene, spec)
            assert ref["rc"] == 0
            u32, f32 = H.gpu_render_level2(r, scene, spec)
            if mode == 0:
                check_frames(scene, u32, f32, ref)
            else:
                assert H.first_mismatch(u32, ref["out"]) is None, scene.name
    finally:
        r.close()


@pytest.mark.parametrize("tile,radix", [(16, 0), (32, 3), (64, 5), (128, 2)])
def test_rule_scenes_tiles_and_radix(tile, radix):
    spec = G.default_spec(max_tile_size=tile, sort_radix_bits=radix)
    r = G.Rasterizer(spec)
    try:
        for scene in rule_scenes() + [S.uniform_512(300, 256, 7)]:
            say(f"{scene.name} tile {tile} radix {radix}")
            ref = H.oracle_render(scene, spec)
            u32, f32 = H.gpu_render_level2(r, scene, spec)
            check_frames(scene, u32, f32, ref)
    finally:
        r.close()


def test_device_serialiser_marks_reversed_strands(rasterizer):
    """The device serialiser writes the reversed-strand flag of nonzero shapes
    exactly as the host serialiser does, and the frame binned on the device
    equals the oracle's."""
    import torch
    for scene in rule_scenes() + [S.uniform_512()]:
        say(f"device serialiser {scene.name}")
        rasterizer.scene_upload_outlines(scene)
        dev = rasterizer.scene_geometry()
        host = G.Geometry(scene)
        assert np.array_equal(dev["geo"], np.asarray(host.geo, np.uint8))
        assert np.array_equal(dev["shapes"], np.asarray(host.shapes, np.uint8))
        flags = 0
        heap = np.asarray(host.geo, np.uint8)
        sh = np.asarray(host.shapes, np.uint8).view(np.uint32).reshape(-1, 4)
        for rec in sh:
            off = 16 * int(rec[2])
            for _ in range(int(rec[3])):
                hd = heap[off:off + 4].view(np.uint16)
                flags |= int(hd[1])
                off += 8 * int(hd[0])
        if scene.name in ("rects", "scribbles", "uniform_512"):
            assert flags == 1, "no reversed strand was marked"
        rasterizer.scene_bin()
        out = torch.zeros(scene.width * scene.height, dtype=torch.int32, device="cuda")
        rasterizer.set_stream(torch.cuda.current_stream().cuda_stream)
        rasterizer.frame_render(out.data_ptr())
        rasterizer.frame_check()
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(np.uint32).reshape(scene.height, scene.width)
        ref = H.oracle_render(scene, want_f64=False)["out"]
        assert H.first_mismatch(got, ref) is None, scene.name


def test_batch_with_nonzero_scenes():
    """Nonzero scenes in one persistent launch (host-built and device-built
    batches, direct and graph replays)."""
    import torch
    r = G.Rasterizer()
    r.set_stream(torch.cuda.current_stream().cuda_stream)
    scenes = rule_scenes() + [S.uniform_512(400, 320, 12)]
    refs = [H.oracle_render(sc, want_f64=False)["out"] for sc in scenes]
    try:
        b = G.Batch(r, [(sc,) + G.prepare(sc)[1:] for sc in scenes])
        d = G.Batch.from_outlines(r, scenes)
        for which, bb in (("host-built", b), ("device-built", d)):
            for graph in (False, True, True):
                say(f"batch {which} graph={graph}")
                bb.render(use_graph=graph)
                bb.check()
                for i, ref in enumerate(refs):
                    assert H.first_mismatch(bb.frame(i), ref) is None, (which, scenes[i].name)
        b.close()
        d.close()
    finally:
        r.close()


def test_adjoint_rejects_nonzero():
    """The adjoint differentiates the crossing-parity fill only: a nonzero
    shape is refused at upload with an error, not rendered wrongly."""
    scene = S.build_scene(32, 32, [(S.nonzero(S.solid_tag(0)), [S.rectangle_pairs(4.5, 4.5, 20.5, 20.5)])],
                          [[0.5, 0.5, 0.5, 1.0]])
    r = G.Rasterizer()
    try:
        with pytest.raises(G.GudniError) as ei:
            G.Adjoint(r, [(scene, np.zeros((32, 32, 4), np.float32))])
        assert ei.value.code == -1 and "nonzero" in str(ei.value).lower()
    finally:
        r.close()
