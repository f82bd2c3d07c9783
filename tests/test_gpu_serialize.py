"""The serialiser on the device: strand heap, Shape records, boxes and strand
extents built from the onShape outlines by gudni_scene_upload_outlines equal
the host serialiser's byte for byte (the host one is pinned by the oracle's
strand round trip and knob-split tests), and the frame rendered from them
equals the host path's and the oracle's."""
import math
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import helpers as H
from tests.test_gpu_binning import empty_scene
from tests.test_gpu_parity import comb_scene, compound_scene, textured_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-serialize] {msg}", flush=True)
    sys.stdout.flush()


def awkward_scene(with_nan=True):
    """Cases the serialiser has to get right: outlines of 0, 1 and 2 pairs,
    vertical and horizontal segments, knobs bulging left and right, a long
    x-monotone run (> 16 curves, so several strands, reversed when it runs
    right to left), a run that wraps around the outline's start, shapes off
    the canvas, a shape with an empty outline beside a real one, and (for the
    geometry comparison only) a NaN coordinate."""
    shapes = []
    # long monotone zig-zag: 40 curves left to right, back along a straight edge
    xs = np.linspace(5, 195, 41)
    top = [(x, 20 + 6 * math.sin(i)) for i, x in enumerate(xs)]
    pts = top + [(195, 60), (5, 60)]
    shapes.append((S.solid_tag(0), [S.straight_pairs(pts)]))
    # knobs: control points outside the end points' x range, both sides
    knob = np.array([[30, 80, 5, 95, 30, 110], [30, 110, 60, 120, 80, 110],
                     [80, 110, 110, 95, 80, 80], [80, 80, 55, 70, 30, 80]], np.float32)
    shapes.append((S.solid_tag(1), [knob[:, :4].copy()]))
    # verticals and horizontals only
    shapes.append((S.solid_tag(2), [S.rectangle_pairs(120, 70, 150, 130)]))
    # degenerate outlines: 0, 1 and 2 pairs, and an empty outline next to a square
    shapes.append((S.solid_tag(3), [np.zeros((0, 4), np.float32)]))
    shapes.append((S.solid_tag(3), [np.array([[10, 150, 20, 160]], np.float32)]))
    shapes.append((S.solid_tag(3), [np.array([[10, 150, 30, 140], [40, 170, 20, 180]], np.float32)]))
    shapes.append((S.solid_tag(4), [np.zeros((0, 4), np.float32), S.rectangle_pairs(160, 140, 190, 170)]))
    # off the canvas on every side
    shapes.append((S.solid_tag(1), [S.rectangle_pairs(-50, 10, -10, 40)]))
    shapes.append((S.solid_tag(1), [S.rectangle_pairs(210, 10, 260, 40)]))
    shapes.append((S.solid_tag(1), [S.rectangle_pairs(10, -60, 40, -5)]))
    shapes.append((S.solid_tag(1), [S.rectangle_pairs(10, 200, 40, 230)]))
    # a run that wraps: the outline starts in the middle of a rising run
    wrap = S.straight_pairs([(100, 150), (140, 185), (60, 190), (70, 170), (90, 160)])
    shapes.append((S.solid_tag(2), [wrap]))
    # a NaN y (kept: its box still meets the canvas; its y extent is unknown)
    if with_nan:
        nan = S.rectangle_pairs(60, 130, 90, 145)
        nan[2, 1] = np.nan
        shapes.append((S.solid_tag(4), [nan]))
    subs = [[1, 0, 0, 0.8], [0, 1, 0, 0.7], [0, 0, 1, 0.9], [1, 1, 0, 1], [0, 1, 1, 0.5]]
    return S.build_scene(200, 200, shapes, subs, name="awkward")


SCENES = {
    "tinySquare": H.tiny_square,
    "compound": compound_scene,
    "textured": textured_scene,
    "comb": comb_scene,
    "empty": empty_scene,
    "awkward": awkward_scene,
    "awkward_finite": lambda: awkward_scene(with_nan=False),
    "uniform_512": lambda: S.uniform_512(),
    "batch_gradients": lambda: S.batch_scene(640, 5000, 3),
    "blobs_1024": lambda: S.blobs_scaled(1024, 12500, 7),
    "adjoint_1024": lambda: S.adjoint_1024(),
}


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def compare_geometry(r, scene):
    geo = G.Geometry(scene)
    r.scene_upload(geo, scene)
    host = r.scene_geometry()
    r.scene_upload_outlines(scene)
    dev = r.scene_geometry()
    assert dev["shapes"].size == geo.shapes.size, f"{scene.name}: {dev['shapes'].size // 16} shapes vs {geo.n_shapes}"
    assert dev["geo"].size == geo.geo.size, f"{scene.name}: heap {dev['geo'].size} vs {geo.geo.size} bytes"
    if geo.geo.size:
        bad = np.flatnonzero(dev["geo"] != geo.geo)
        assert bad.size == 0, f"{scene.name}: heap differs first at byte {int(bad[0])}"
    np.testing.assert_array_equal(dev["shapes"], geo.shapes, err_msg="shape records")
    assert same_bits(dev["bbox"], geo.bbox.reshape(-1, 4)), "boxes"
    assert same_bits(dev["bounds"], host["bounds"]), "strand extents"
    return geo


@pytest.mark.parametrize("name", list(SCENES))
def test_device_serialiser_equals_host(rasterizer, name):
    scene = SCENES[name]()
    say(f"{name} {scene.width}x{scene.height} {scene.n_shapes} shapes {scene.n_curves} curves")
    compare_geometry(rasterizer, scene)


def test_device_serialiser_full_blobs_4096(rasterizer):
    scene = S.blobs_4096()
    say("blobs 4096^2 geometry")
    compare_geometry(rasterizer, scene)


@pytest.mark.parametrize("name", ["awkward_finite", "textured", "blobs_1024"])
def test_frame_from_outlines(rasterizer, name):
    """Outlines -> device serialiser -> device binning -> render gives the
    host path's frame and the oracle's."""
    import torch
    scene = SCENES[name]()
    say(f"{name}: host path")
    geo = G.Geometry(scene)
    rasterizer.scene_upload(geo, scene)
    rasterizer.scene_bin()
    n = scene.width * scene.height
    a = torch.zeros(n, dtype=torch.int32, device="cuda")
    rasterizer.frame_render(a.data_ptr())
    rasterizer.frame_check()
    say(f"{name}: device serialiser")
    rasterizer.scene_upload_outlines(scene)
    rasterizer.scene_bin()
    b = torch.zeros(n, dtype=torch.int32, device="cuda")
    rasterizer.frame_render(b.data_ptr())
    rasterizer.frame_check()
    assert torch.equal(a, b)
    ref = H.oracle_render(scene)
    assert ref["rc"] == 0
    got = b.cpu().numpy().view(np.uint32).reshape(np.shape(ref["out"]))
    assert H.first_mismatch(got, ref["out"]) is None


def test_bad_substance_is_an_error(rasterizer):
    scene = S.build_scene(64, 64, [(S.solid_tag(5), [S.rectangle_pairs(4, 4, 20, 20)])], [[1, 0, 0, 1]],
                          name="bad-substance")
    with pytest.raises(G.GudniError) as ei:
        rasterizer.scene_upload_outlines(scene)
    assert "substance" in str(ei.value)
    # the context still works afterwards
    rasterizer.scene_upload_outlines(H.tiny_square())
    rasterizer.scene_bin()
