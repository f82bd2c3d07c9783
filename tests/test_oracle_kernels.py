"""Pins the oracle's kernel restatement (generate / sort / render) against
facts fixed by the mathematics of the reference's threshold model, each
computed by an algorithm independent of the oracle:

  * axis-aligned rectangles whose vertical edges are >= MINCROP tall give the
    exact pixel-overlap area (closed form);
  * convex polygons whose vertices have integer x give the exact polygon /
    pixel intersection area (Sutherland-Hodgman clipping in FP64): every
    threshold is then a persistent bridge across its column;
  * stacked translucent full-cover shapes give the premultiplied closed form
    sum_i c_i a_i prod_{j<i} (1 - a_j) + bg prod (1 - a_j), first shape on top;
  * the sorted queue equals numpy's stable lexsort of the generated queue by
    (top, topX, invertedSlope);
  * nearest-texel sampling equals a numpy gather at rtz((x, y)/scale - t);
  * the 127-shape cap keeps the 127 bottom-most shapes of a column.
"""
import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from oracle import oracle as O
from tests import helpers as H

TOL = 1e-6


def rect_cover(W, Hh, x0, y0, x1, y1):
    xs = np.arange(W, dtype=np.float64)
    ys = np.arange(Hh, dtype=np.float64)
    ox = np.clip(np.minimum(xs + 1, x1) - np.maximum(xs, x0), 0, None)
    oy = np.clip(np.minimum(ys + 1, y1) - np.maximum(ys, y0), 0, None)
    return oy[:, None] * ox[None, :]


def clip_area(poly, x0, y0, x1, y1):
    """Area of polygon n [x0,x1]x[y0,y1] by Sutherland-Hodgman, FP64."""
    def clip(pts, inside, cross):
        out = []
        for i in range(len(pts)):
            a, b = pts[i - 1], pts[i]
            ia, ib = inside(a), inside(b)
            if ib:
                if not ia:
                    out.append(cross(a, b))
                out.append(b)
            elif ia:
                out.append(cross(a, b))
        return out

    def cx(a, b, x):
        t = (x - a[0]) / (b[0] - a[0])
        return (x, a[1] + t * (b[1] - a[1]))

    def cy(a, b, y):
        t = (y - a[1]) / (b[1] - a[1])
        return (a[0] + t * (b[0] - a[0]), y)

    pts = [tuple(map(float, p)) for p in poly]
    pts = clip(pts, lambda p: p[0] >= x0, lambda a, b: cx(a, b, x0))
    if pts:
        pts = clip(pts, lambda p: p[0] <= x1, lambda a, b: cx(a, b, x1))
    if pts:
        pts = clip(pts, lambda p: p[1] >= y0, lambda a, b: cy(a, b, y0))
    if pts:
        pts = clip(pts, lambda p: p[1] <= y1, lambda a, b: cy(a, b, y1))
    if len(pts) < 3:
        return 0.0
    p = np.array(pts)
    return 0.5 * abs(np.dot(p[:, 0], np.roll(p[:, 1], -1)) - np.dot(p[:, 1], np.roll(p[:, 0], -1)))


def coverage_of_black(res):
    """Black opaque shape over white: coverage = 1 - red."""
    return 1.0 - res["f64"][:, :, 0]


@pytest.mark.parametrize("seed", range(12))
def test_rectangles_exact(seed):
    rng = np.random.default_rng(seed)
    W = Hh = int(rng.choice([16, 64, 300]))
    x0, x1 = sorted(rng.uniform(0, W, 2))
    y0 = rng.uniform(0, Hh - 2)
    y1 = y0 + rng.uniform(1.0, Hh - y0)
    x0, y0, x1, y1 = [float(np.float32(v)) for v in (x0, y0, x1, y1)]
    scene = S.build_scene(W, Hh, [(S.solid_tag(0), [S.rectangle_pairs(x0, y0, x1, y1)])],
                          [[0, 0, 0, 1]], name="rect")
    res = H.oracle_render(scene)
    assert res["rc"] == 0
    np.testing.assert_allclose(coverage_of_black(res), rect_cover(W, Hh, x0, y0, x1, y1), atol=TOL)


def convex_polygon(rng, W, Hh, n):
    """Random convex polygon with integer vertex x and no vertical edge."""
    cx, cy = rng.uniform(0.3 * W, 0.7 * W), rng.uniform(0.3 * Hh, 0.7 * Hh)
    r = rng.uniform(0.15, 0.3) * min(W, Hh)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    pts = [(np.round(cx + r * np.cos(a)), cy + r * np.sin(a)) for a in ang]
    # convex hull (monotone chain) keeps it convex after rounding x
    pts = sorted(set((float(x), float(np.float32(y))) for x, y in pts))
    if len(pts) < 3:
        return None

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])
    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    hull = lower[:-1] + upper[:-1]
    xs = [p[0] for p in hull]
    for i in range(len(hull)):
        if xs[i] == xs[i - 1]:
            return None          # vertical edge
    return hull if len(hull) >= 3 else None


@pytest.mark.parametrize("seed", range(10))
def test_convex_polygons_exact_area(seed):
    rng = np.random.default_rng(100 + seed)
    W = Hh = int(rng.choice([32, 96, 260]))
    poly = None
    while poly is None:
        poly = convex_polygon(rng, W, Hh, int(rng.integers(3, 9)))
    scene = S.build_scene(W, Hh, [(S.solid_tag(0), [S.straight_pairs(poly)])], [[0, 0, 0, 1]],
                          name="convex")
    res = H.oracle_render(scene)
    assert res["rc"] == 0
    cov = coverage_of_black(res)
    exact = np.array([[clip_area(poly, x, y, x + 1, y + 1) for x in range(W)] for y in range(Hh)])
    np.testing.assert_allclose(cov, exact, atol=1e-5)


def test_porter_duff_stack_closed_form():
    """Six translucent full-canvas rectangles; scene order = front to back."""
    rng = np.random.default_rng(7)
    n = 6
    cols = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    shapes = [(S.solid_tag(i), [S.rectangle_pairs(-1, -1, 40, 40)]) for i in range(n)]
    bg = np.array([0.2, 0.3, 0.4, 1.0])
    scene = S.build_scene(32, 32, shapes, cols, background=tuple(bg), name="stack")
    res = H.oracle_render(scene)
    c = cols.astype(np.float64)
    trans = np.concatenate([[1.0], np.cumprod(1 - c[:, 3])])
    expect = (c[:, :3] * (c[:, 3] * trans[:-1])[:, None]).sum(0) + bg[:3] * trans[-1]
    got = res["f64"][:, :, :3].reshape(-1, 3)
    assert np.abs(got - expect).max() < 1e-5


def test_sorted_queue_is_stable_lexsort():
    """The sort stage equals numpy's stable sort by (top, topX, invSlope)."""
    scene = S.uniform_512(300, 256, 21)
    spec = O.default_spec()
    g = O.Geometry(scene)
    j = O.JobList(spec, g, scene.width, scene.height)
    rc, _, d = O.raster_job_dump(scene, spec, g, j, 0, cap=96)
    assert rc == 0
    checked = 0
    for slot in np.nonzero(d["qlen"] > 1)[0]:
        n = d["qlen"][slot]
        assert n <= 96
        thr, hdr = d["pre_thr"][slot, :n], d["pre_hdr"][slot, :n]
        pos = (hdr & 0x80000000) != 0
        top, bot, left, right = (thr[:, k].astype(np.float32) for k in range(4))
        topx = np.where(pos, left, right)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.where(top == bot, np.float32(np.finfo(np.float32).max),
                           ((right - left) / (bot - top)).astype(np.float32)
                           * np.where(pos, np.float32(1), np.float32(-1)))
        # canonicalise -0 so it ties with +0 like the float compare does
        keys = [np.where(k == 0, np.float32(0), k) for k in (inv, topx, top)]
        order = np.lexsort(keys)          # stable: ties keep queue order
        np.testing.assert_array_equal(d["thr"][slot, :n].view(np.uint32), thr[order].view(np.uint32))
        np.testing.assert_array_equal(d["hdr"][slot, :n], hdr[order])
        checked += 1
    assert checked > 100


def test_nearest_texture_sampling():
    rng = S.SplitMix64(5)
    tex = (rng.next(16 * 16) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.uint8)
    scale, tx, ty = 0.5, -2.0, -1.0
    scene = S.build_scene(24, 24, [(S.picture_tag(0), [S.rectangle_pairs(0, 0, 24, 24)])],
                          [S.picture_substance(0)], tex, [(tx, ty, 16, 16, 0, scale)],
                          background=(0, 0, 0, 1), name="tex")
    res = H.oracle_render(scene)
    t = tex.reshape(16, 16, 4).astype(np.float64) / 255.0
    for y in range(24):
        for x in range(24):
            u = np.trunc(np.float32(x) / np.float32(scale) - np.float32(tx))
            v = np.trunc(np.float32(y) / np.float32(scale) - np.float32(ty))
            if 0 <= u < 16 and 0 <= v < 16:
                texel = t[int(v), int(u)]
                a = texel[3]
                expect = texel[:3] * a       # over opaque black
            else:
                expect = np.zeros(3)
            assert np.abs(res["f64"][y, x, :3] - expect).max() < 1e-5, (x, y)


def test_shape_cap_keeps_bottom_most():
    """200 opaque full-cover shapes: a column keeps the first 127 of its tile
    list (newest first = bottom-most), so shape 200-127 = 73 is on top."""
    n = 200
    cols = np.zeros((n, 4), np.float32)
    cols[:, 0] = np.arange(n) / 255.0
    cols[:, 3] = 1.0
    shapes = [(S.solid_tag(i), [S.rectangle_pairs(-1, -1, 20, 20)]) for i in range(n)]
    scene = S.build_scene(16, 16, shapes, cols, name="cap")
    res = H.oracle_render(scene)
    assert np.allclose(res["f64"][:, :, 0], 73 / 255.0, atol=1e-6)


def test_queue_overflow_reported():
    scene = S.uniform_512(100, 64, 3)
    res = H.oracle_render(scene, G.default_spec(max_thresholds=2))
    assert res["rc"] == -1


def test_accurate_coverage_within_tolerance():
    """Accurate mode stays within 1e-5 of the FP64 integration on the
    reference's workloads (the 512^2 uniform scene and a 1024^2 version of the 4096^2 blob scene)."""
    for scene in (S.uniform_512(300, 256, 4), S.blobs_scaled(512, 3000, 8)):
        res = H.oracle_render(scene)
        assert np.abs(res["f32"] - res["f64"]).max() < 1e-5


def test_reference_mode_differs_only_in_section_width():
    """Both coverage modes produce the same queue dumps; frames differ by at
    most one 8-bit step."""
    scene = S.blobs_scaled(512, 3000, 8)
    a = H.oracle_render(scene, G.default_spec(coverage_mode=0))
    b = H.oracle_render(scene, G.default_spec(coverage_mode=1))
    da = a["out"].view(np.uint8).astype(int)
    db = b["out"].view(np.uint8).astype(int)
    assert np.abs(da - db).max() <= 1


def test_bilinear_texture_sampling():
    """Bilinear pictures: each pixel blends the four texels around
    p/scale - translate - 0.5, texels outside the picture transparent."""
    rng = S.SplitMix64(7)
    tex = (rng.next(12 * 12) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.uint8)
    scale, tx, ty = 0.75, -1.5, -2.25
    scene = S.build_scene(20, 20, [(S.picture_tag(0), [S.rectangle_pairs(0, 0, 20, 20)])],
                          [S.picture_substance(0)], tex, [(tx, ty, 12, 12, 0, scale)],
                          background=(0, 0, 0, 1), name="bilinear")
    spec = G.default_spec(texture_filter=1)
    res = H.oracle_render(scene, spec)
    t = tex.reshape(12, 12, 4).astype(np.float64) / 255.0

    def at(i, j):
        return t[j, i] if 0 <= i < 12 and 0 <= j < 12 else np.zeros(4)
    for y in range(20):
        for x in range(20):
            sx = float(np.float32(np.float32(x) / np.float32(scale)) - np.float32(tx)) - 0.5
            sy = float(np.float32(np.float32(y) / np.float32(scale)) - np.float32(ty)) - 0.5
            ix, iy = int(np.floor(sx)), int(np.floor(sy))
            wx, wy = sx - ix, sy - iy
            c = (at(ix, iy) * (1 - wx) + at(ix + 1, iy) * wx) * (1 - wy) + \
                (at(ix, iy + 1) * (1 - wx) + at(ix + 1, iy + 1) * wx) * wy
            expect = c[:3] * c[3]            # over opaque black
            assert np.abs(res["f64"][y, x, :3] - expect).max() < 1e-5, (x, y)


@pytest.mark.parametrize("kind", [0, 1])
def test_gradient_substances(kind):
    """Linear and radial gradients over a covering rectangle: every pixel is
    the piecewise-linear stop colour at t of its integer position."""
    offs = np.array([0.0, 0.3, 1.0])
    cols = np.array([[1.0, 0.0, 0.0, 1.0], [0.0, 1.0, 0.0, 0.5], [0.0, 0.0, 1.0, 1.0]])
    if kind == 0:
        g = S.gradient_record(0, (2.0, 3.0), (14.0, 9.0), offs, cols)
    else:
        g = S.gradient_record(1, (8.0, 7.0), (9.0, 0.0), offs, cols)
    scene = S.build_scene(16, 16, [(S.gradient_tag(0), [S.rectangle_pairs(0, 0, 16, 16)])],
                          [S.gradient_substance(0)], background=(1, 1, 1, 1), name="grad",
                          gradients=[g])
    res = H.oracle_render(scene)
    for y in range(16):
        for x in range(16):
            if kind == 0:
                d = np.array([12.0, 6.0])
                t = ((x - 2.0) * d[0] + (y - 3.0) * d[1]) / (d @ d)
            else:
                t = np.hypot(x - 8.0, y - 7.0) / 9.0
            t = min(max(t, 0.0), 1.0)
            i = int(np.searchsorted(offs, t, side="left"))
            if i == 0:
                c = cols[0]
            else:
                f = (t - offs[i - 1]) / (offs[i] - offs[i - 1])
                c = cols[i - 1] + (cols[i] - cols[i - 1]) * f
            expect = c[:3] * c[3] + (1 - c[3]) * 1.0      # over white
            assert np.abs(res["f64"][y, x, :3] - expect).max() < 1e-5, (x, y, t)
