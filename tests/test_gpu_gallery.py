"""Edge cases from the reference's scene gallery (benchmarks/GudniTests.hs:65-98,
100-126, 247-420): the reference shows them in a window and asserts nothing, so
their geometry is restated here and the GPU frames are checked against the
oracle in both coverage modes (BGRA bit for bit, pre-quantisation colour within
1e-5 of the FP64 integration)."""
import math
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import helpers as H
from tests.test_gpu_parity import check_frames

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-gallery] {msg}", flush=True)
    sys.stdout.flush()


def xform(points, scale=1.0, angle=0.0, tx=0.0, ty=0.0):
    """Scale, then rotate about the origin, then translate (the order the
    gallery's tScale / tRotate / tTranslate compose in)."""
    c, s = math.cos(angle), math.sin(angle)
    out = []
    for x, y in points:
        x, y = x * scale, y * scale
        out.append((x * c - y * s + tx, x * s + y * c + ty))
    return out


def rect(x0, y0, w, h):
    return [(x0, y0), (x0 + w, y0), (x0 + w, y0 + h), (x0, y0 + h)]


def poly(points, **kw):
    return S.straight_pairs(xform(points, **kw))


def simple_knob():
    """A curve whose control point bulges right of both end points
    (:247-256): split at its vertical tangent before serialising."""
    pairs = np.array([[0, 0, 1, 1], [0, 2, 0, 1]], np.float32) * 100 + 100
    return S.build_scene(320, 320, [(S.solid_tag(0), [pairs])], [[0.25, 0.25, 0.25, 1.0]],
                         name="simpleKnob")


def hour_glass():
    """Self-crossing edges (:258-267)."""
    return S.build_scene(16, 16, [(S.solid_tag(0), [poly([(0, 0), (1, 1), (1, 0), (0, 1)], scale=8)])],
                         [[0.25, 0.25, 0.25, 1.0]], name="hourGlass")


def six_point_rectangle():
    """Collinear points on horizontal edges (:352-358)."""
    pts = [(0, 0), (1, 0), (2, 0), (2, 1), (1, 1), (0, 1)]
    return S.build_scene(8, 8, [(S.solid_tag(0), [poly(pts)])], [[0.3, 0.3, 0.3, 1.0]],
                         name="sixPointRectangle")


def open_square():
    """Subtraction with transparency (:294-299)."""
    shapes = [(S.solid_tag(0), [poly(rect(0, 0, 5, 5))]),
              (S.solid_tag(0, subtract=True), [poly(rect(1, 1, 3, 3))])]
    return S.build_scene(8, 8, shapes, [[1.0, 0.5, 0.0, 0.5]], name="openSquare")


def open_square_overlap2():
    """Two subtracted squares overlapping, translucent over opaque (:301-307)."""
    shapes = []
    for k, (t, sid) in enumerate([(0, 0), (8, 1)]):
        shapes.append((S.solid_tag(sid), [poly(rect(0, 0, 16, 16), scale=20, tx=20 * t, ty=20 * t)]))
        shapes.append((S.solid_tag(sid, subtract=True), [poly(rect(4, 4, 8, 8), scale=20, tx=20 * t, ty=20 * t)]))
    return S.build_scene(512, 512, shapes, [[0, 0, 1, 0.25], [1, 0.5, 0, 1.0]], name="openSquareOverlap2")


def open_square_overlap3(angle=0.7):
    """Three rotated subtracted squares, all translucent (:309-318)."""
    shapes = []
    for k, (t, a) in enumerate([(0, angle), (4, angle / 2), (8, angle / 3)]):
        o = dict(scale=50, angle=a, tx=400 + 50 * t, ty=400 + 50 * t)
        shapes.append((S.solid_tag(k), [poly(rect(0, 0, 16, 16), **o)]))
        shapes.append((S.solid_tag(k, subtract=True), [poly(rect(4, 4, 8, 8), **o)]))
    return S.build_scene(1024, 1024, shapes, [[0, 0, 1, 0.5], [1, 0.5, 0, 0.5], [0, 1, 0, 0.5]],
                         name="openSquareOverlap3")


def concentric_squares3():
    """Abutting edges of nested compounds (:320-333)."""
    shapes = [(S.solid_tag(0), [poly(rect(0, 0, 10, 10))]),
              (S.solid_tag(0, subtract=True), [poly(rect(2, 2, 6, 6))]),
              (S.solid_tag(1), [poly(rect(2, 2, 6, 6))]),
              (S.solid_tag(1, subtract=True), [poly(rect(4, 4, 2, 2))]),
              (S.solid_tag(2), [poly(rect(4, 4, 2, 2))])]
    return S.build_scene(12, 12, shapes, [[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]],
                         name="concentricSquares3")


def two_brackets():
    """Sub-pixel slivers leaving the canvas on the left (:398-406)."""
    a = [(-0.2, 0.0), (1.2, 1.4), (1.2, 1.6), (-0.2, 0.2)]
    b = [(-0.2, 0.4), (1.2, 1.8), (1.2, 2.0), (-0.2, 0.6)]
    return S.build_scene(4, 4, [(S.solid_tag(0), [poly(a)]), (S.solid_tag(1), [poly(b)])],
                         [[1, 1, 1, 1], [1, 1, 1, 1]], background=(0, 0, 0, 1), name="twoBrackets")


def tall_rectangle():
    """A half-pixel-wide rectangle, rotated 3 degrees, crossing many tiles
    (tallRectangle, :384-389), on a non-square canvas."""
    return S.build_scene(128, 1024, [(S.solid_tag(0), [poly(rect(0, 0, 0.5, 2000), angle=-math.radians(3),
                                                         tx=0, ty=1)])],
                         [[1, 1, 1, 1]], background=(0, 0, 0, 1), name="tallRectangle")


def max_shapes(n=2001):
    """Very wide one-pixel rectangles stacked half a pixel apart, red, green
    and blue in turn (maxShapeTest, :115-124): many shapes per tile, the
    tile tree down to its floor."""
    cols = [[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]]
    shapes = [(S.solid_tag(i % 3), [poly(rect(-10, 0.5 * i, 10000, 1))]) for i in range(n)]
    return S.build_scene(512, 512, shapes, cols, name="maxShapeTest")


def max_thresholds(n=400):
    """Long thin rotated rectangles a quarter pixel apart (maxThresholdTest,
    :100-111): many records per column."""
    shapes = [(S.solid_tag(0), [poly(rect(0, 0, 1000, 1), scale=0.1, angle=0.2, tx=-1, ty=1 + 0.25 * i)])
              for i in range(n)]
    return S.build_scene(128, 128, shapes, [[0.6, 0.4, 0.8, 1.0]], name="maxThresholdTest")


def subtract_diamond():
    """A subtraction whose subtracted shape only touches a corner (:408-412)."""
    sq = rect(0, 0, 1, 1)
    shapes = [(S.solid_tag(0), [poly(sq, scale=4, angle=math.pi / 4, tx=6, ty=1)]),
              (S.solid_tag(0, subtract=True), [poly(rect(1, 1, 1, 1), scale=4, angle=math.pi / 4, tx=6, ty=1)])]
    return S.build_scene(12, 12, shapes, [[1, 1, 1, 1]], background=(0, 0, 0, 1), name="subtractDiamond")


SCENES = {f.__name__: f for f in (simple_knob, hour_glass, six_point_rectangle, open_square,
                                  open_square_overlap2, open_square_overlap3, concentric_squares3,
                                  two_brackets, tall_rectangle, max_shapes, max_thresholds,
                                  subtract_diamond)}


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name", list(SCENES))
def test_gallery_scene(name, mode):
    scene = SCENES[name]()
    spec = G.default_spec(coverage_mode=mode)
    say(f"{name} mode {mode}: {scene.width}x{scene.height}, {scene.n_shapes} shapes")
    ref = H.oracle_render(scene, spec)
    assert ref["rc"] == 0
    r = G.Rasterizer(spec)
    try:
        u32, f32 = H.gpu_render_level2(r, scene, spec)
    finally:
        r.close()
    check_frames(scene, u32, f32, ref)
