"""Pins the oracle's nonzero fill rule (an extension: the reference fills by
crossing parity only, Kernels.cl:298-304, 1515-1519) against facts fixed by
the geometry, each computed without the oracle's sweep:

  * two overlapping rectangles traced the same way round are one region
    under nonzero (union area) and two under even-odd (symmetric
    difference); traced opposite ways round they are the symmetric
    difference under both rules (closed-form pixel overlaps);
  * a pentagram with integer-x vertices: nonzero covers the whole star
    (exact FP64 clip of its 10-gon boundary), even-odd leaves the inner
    pentagon empty (star minus pentagon), away from the five pixels where
    edges cross inside a section;
  * reversing every outline leaves nonzero coverage unchanged;
  * random self-intersecting polygons: every pixel no edge comes near has
    coverage exactly (winding number at its centre != 0), the winding
    number computed by a plain ray-crossing count;
  * a scene of simple closed shapes renders the same bits under both rules.
"""
import numpy as np
import pytest

from gudni_amd import scenes as S
from oracle import oracle as O
from tests.test_oracle_kernels import clip_area, rect_cover

TOL = 1e-6


def cover(scene):
    res = O.render(scene, O.default_spec(), want_f64=True)
    assert res["rc"] == 0
    return 1.0 - res["f64"][:, :, 0], res


def black(tag_fn, outlines, W, Hh):
    return S.build_scene(W, Hh, [(tag_fn(S.solid_tag(0)), outlines)], [[0, 0, 0, 1]])


def rect_pts(x0, y0, x1, y1, reverse=False):
    pts = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    return pts[::-1] if reverse else pts


@pytest.mark.parametrize("reverse", [False, True])
def test_overlapping_rectangles(reverse):
    W = Hh = 48
    a = (4, 5.25, 30, 33.5)
    b = (17, 12.75, 41, 44.0)
    ca, cb = rect_cover(W, Hh, *a), rect_cover(W, Hh, *b)
    ov = rect_cover(W, Hh, max(a[0], b[0]), max(a[1], b[1]), min(a[2], b[2]), min(a[3], b[3]))
    outlines = [S.straight_pairs(rect_pts(*a)), S.straight_pairs(rect_pts(*b, reverse=reverse))]
    nz, _ = cover(black(S.nonzero, outlines, W, Hh))
    eo, _ = cover(black(lambda t: t, outlines, W, Hh))
    sym = ca + cb - 2 * ov
    assert np.abs(eo - sym).max() < TOL
    want = sym if reverse else ca + cb - ov
    assert np.abs(nz - want).max() < TOL
    if not reverse:
        assert nz[20, 20] == pytest.approx(1.0) and eo[20, 20] == pytest.approx(0.0)


def star_points(cx, cy, r):
    """Pentagram vertices (every second point of a regular pentagon), x
    rounded to integers so every edge spans whole columns.  Turned by 0.1 rad
    so that no edge is horizontal: where a horizontal record and a sloped one
    start at the same end point the reference's sweep leaves the horizontal
    one active (verticalAdvance only drops horizontal records at the front of
    the queue, Kernels.cl:1744-1824) and gives the pixel a negative section
    under either fill rule."""
    ang = -np.pi / 2 + 0.1 + np.arange(5) * 4 * np.pi / 5
    return [(float(np.round(cx + r * np.cos(t))), float(cy + r * np.sin(t))) for t in ang]


def seg_intersect(p, q, r, s):
    p, q, r, s = map(np.asarray, (p, q, r, s))
    d1, d2 = q - p, s - r
    den = d1[0] * d2[1] - d1[1] * d2[0]
    if den == 0:
        return None
    t = ((r - p)[0] * d2[1] - (r - p)[1] * d2[0]) / den
    u = ((r - p)[0] * d1[1] - (r - p)[1] * d1[0]) / den
    if 0 < t < 1 and 0 < u < 1:
        return p + t * d1
    return None


def test_pentagram():
    W = Hh = 64
    pts = star_points(32, 33, 27)
    edges = [(pts[i], pts[(i + 1) % 5]) for i in range(5)]
    crossings = []
    for i in range(5):
        for j in range(i + 1, 5):
            x = seg_intersect(*edges[i], *edges[j])
            if x is not None:
                crossings.append(x)
    assert len(crossings) == 5
    # the star's boundary: outer vertex, inner crossing, ... (angular order)
    c = np.array([32.0, 33.0])
    boundary = sorted([np.asarray(p) for p in pts] + crossings,
                      key=lambda v: np.arctan2(v[1] - c[1], v[0] - c[0]))
    inner = sorted(crossings, key=lambda v: np.arctan2(v[1] - c[1], v[0] - c[0]))
    star = np.array([[clip_area(boundary, x, y, x + 1, y + 1) for x in range(W)] for y in range(Hh)])
    pent = np.array([[clip_area(inner, x, y, x + 1, y + 1) for x in range(W)] for y in range(Hh)])
    outline = [S.straight_pairs(pts)]
    nz, _ = cover(black(S.nonzero, outline, W, Hh))
    eo, _ = cover(black(lambda t: t, outline, W, Hh))
    keep = np.ones((Hh, W), bool)
    for x in crossings:   # sections meet a crossing of two records there
        keep[int(np.floor(x[1])) - 1:int(np.floor(x[1])) + 2, int(np.floor(x[0])) - 1:int(np.floor(x[0])) + 2] = False
    assert np.abs(nz - star)[keep].max() < 1e-5
    assert np.abs(eo - (star - pent))[keep].max() < 1e-5
    assert nz[33, 32] == pytest.approx(1.0) and eo[33, 32] == pytest.approx(0.0)


def reverse_outline(pairs):
    """The same closed curve traced backwards: curve i runs on_i -> off_i ->
    on_(i+1), so backwards it is on_(i+1) -> off_i -> on_i."""
    p = np.asarray(pairs, np.float32).reshape(-1, 4)
    on = np.roll(p[:, :2], -1, axis=0)
    return np.concatenate([on, p[:, 2:]], axis=1)[::-1].copy()


def test_reversing_outlines_keeps_nonzero_coverage():
    rng = np.random.default_rng(5)
    W = Hh = 96
    shapes, shapes_rev = [], []
    for i in range(6):
        k = int(rng.integers(5, 10))
        pts = [(float(rng.integers(0, W)), float(rng.uniform(0, Hh))) for _ in range(k)]
        o = S.straight_pairs(pts)
        shapes.append((S.nonzero(S.solid_tag(i)), [o]))
        shapes_rev.append((S.nonzero(S.solid_tag(i)), [reverse_outline(o)]))
    cols = [[0.1 * i, 0.5, 1 - 0.1 * i, 0.7] for i in range(6)]
    a = O.render(S.build_scene(W, Hh, shapes, cols), O.default_spec(), want_f64=True)
    b = O.render(S.build_scene(W, Hh, shapes_rev, cols), O.default_spec(), want_f64=True)
    assert a["rc"] == 0 and b["rc"] == 0
    assert np.abs(a["f64"] - b["f64"]).max() < 1e-5


def winding(px, py, poly):
    """Winding number of (px, py) (y down) by a ray-crossing count to +x:
    an edge crossing the ray upward counts +1, downward -1."""
    w = 0
    n = len(poly)
    for i in range(n):
        (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % n]
        if (y0 <= py) != (y1 <= py):
            x = x0 + (py - y0) * (x1 - x0) / (y1 - y0)
            if x > px:
                w += 1 if y1 < y0 else -1
    return w


def seg_dist(px, py, a, b):
    a, b, p = np.asarray(a, float), np.asarray(b, float), np.array([px, py])
    d = b - a
    t = 0.0 if not d.any() else float(np.clip(np.dot(p - a, d) / np.dot(d, d), 0, 1))
    return float(np.linalg.norm(p - (a + t * d)))


@pytest.mark.parametrize("seed", range(6))
def test_random_polygons_interior_winding(seed):
    rng = np.random.default_rng(100 + seed)
    W = Hh = 40
    k = int(rng.integers(5, 10))
    poly = [(float(rng.uniform(2, W - 2)), float(rng.uniform(2, Hh - 2))) for _ in range(k)]
    outline = [S.straight_pairs(poly)]
    nz, _ = cover(black(S.nonzero, outline, W, Hh))
    eo, _ = cover(black(lambda t: t, outline, W, Hh))
    polyf = [tuple(map(float, np.float32(p))) for p in poly]
    checked = 0
    for y in range(Hh):
        for x in range(W):
            cx, cy = x + 0.5, y + 0.5
            if min(seg_dist(cx, cy, polyf[i], polyf[(i + 1) % k]) for i in range(k)) <= 0.75:
                continue
            w = winding(cx, cy, polyf)
            checked += 1
            assert nz[y, x] == pytest.approx(1.0 if w != 0 else 0.0, abs=1e-9), (x, y, w)
            assert eo[y, x] == pytest.approx(1.0 if w % 2 else 0.0, abs=1e-9), (x, y, w)
    assert checked > 200


def test_simple_shapes_render_the_same_under_both_rules():
    """Blobs are simple closed curves: winding numbers are 0 or +-1, so the
    nonzero frame equals the even-odd one bit for bit."""
    sc = S.blobs_scaled(256, 300, seed=4, glyph_share=0.0)
    sc_nz = S.Scene(**{**sc.__dict__, "tags": np.array([S.nonzero(t) for t in sc.tags], np.uint64)})
    a = O.render(sc, O.default_spec())
    b = O.render(sc_nz, O.default_spec())
    assert a["rc"] == 0 and b["rc"] == 0
    assert np.array_equal(a["out"], b["out"])
