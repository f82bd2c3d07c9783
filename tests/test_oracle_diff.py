"""The FP64 forward used for the adjoint's finite differences (oracle/
gudni_diff.c) against the oracle's FP32 raster path and its FP64 section
integration: the same pixels, so the finite differences differentiate the
function the raster path computes."""
import numpy as np

from gudni_amd import scenes as S
from oracle import oracle as O
from tests import helpers as H
from tests import adjoint_cases as A


def fp64_image(scene):
    target = np.zeros((scene.height, scene.width, 4), np.float32)
    _, img, _ = O.diff_loss(scene, scene.pairs.reshape(-1, 4), np.asarray(scene.substances, np.float64),
                            target, (0, 0, scene.width, scene.height), want_image=True)
    return img


def polygon_scene(n=120, size=192, seed=3):
    rng = np.random.default_rng(seed)
    shapes, subs = [], []
    for i in range(n):
        k = int(rng.integers(3, 9))
        c = rng.uniform(0, size, 2)
        r = rng.uniform(4, 40)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        pts = [(float(c[0] + r * np.cos(a)), float(c[1] + r * np.sin(a))) for a in ang]
        shapes.append((S.solid_tag(i), [S.straight_pairs(pts)]))
        subs.append([float(v) for v in rng.uniform(0, 1, 4)])
    return S.build_scene(size, size, shapes, subs, name="polygons")


def test_fp64_forward_matches_oracle_on_polygons():
    """Straight edges (no knobs): only FP32 rounding separates the two."""
    sc = polygon_scene()
    ref = H.oracle_render(sc, want_f64=True)
    d = np.abs(fp64_image(sc) - ref["f64"].reshape(sc.height, sc.width, 4)).max(axis=2)
    assert np.median(d) < 1e-7
    assert (d > 1e-5).mean() < 1e-3, f"{(d > 1e-5).sum()} pixels differ by more than 1e-5"


def test_fp64_forward_matches_oracle_on_blobs():
    """Blobs with knob splits: the FP64 forward splits at the exact tangent,
    the serialiser at its FP32 bisection (within 1e-4 in t)."""
    sc = S.batch_scene(192, 200, 5, gradient_share=0)
    ref = H.oracle_render(sc, want_f64=True)
    d = np.abs(fp64_image(sc) - ref["f64"].reshape(sc.height, sc.width, 4)).max(axis=2)
    assert np.median(d) < 1e-7
    assert (d > 1e-5).mean() < 2e-2
    assert d.max() < 2e-2


def test_window_equals_whole_frame():
    """Cutting columns into segments changes nothing (tiles only do that)."""
    sc = polygon_scene(60, 128, 9)
    full = fp64_image(sc)
    target = np.zeros((sc.height, sc.width, 4), np.float32)
    _, win, _ = O.diff_loss(sc, sc.pairs.reshape(-1, 4), np.asarray(sc.substances, np.float64), target,
                            (30, 41, 97, 100), want_image=True)
    np.testing.assert_allclose(win, full[41:100, 30:97], atol=1e-12)


def test_loss_is_sum_of_squares_and_signature_detects_flips():
    sc = polygon_scene(40, 96, 11)
    target = np.random.default_rng(0).uniform(0, 1, (sc.height, sc.width, 4)).astype(np.float32)
    pairs = sc.pairs.reshape(-1, 4).astype(np.float64)
    rgba = np.asarray(sc.substances, np.float64)
    loss, img, sig = O.diff_loss(sc, pairs, rgba, target, (0, 0, 96, 96), want_image=True)
    assert abs(loss - ((img - target) ** 2).sum()) < 1e-9 * max(1.0, loss)
    # a colour change keeps every decision; moving a vertex by pixels does not
    _, _, sig_c = O.diff_loss(sc, pairs, rgba + 0.01, target, (0, 0, 96, 96))
    assert sig_c == sig
    moved = pairs.copy()
    moved[0, :2] += 5.0
    _, _, sig_m = O.diff_loss(sc, moved, rgba, target, (0, 0, 96, 96))
    assert sig_m != sig


def test_colour_gradient_closed_form():
    """One opaque-background pixel fully inside one translucent square: the
    pixel is a*c + (1-a)*bg, so dL/dc = 2 (pix - t) a exactly; the central
    difference of the FP64 forward reproduces it."""
    sq = S.build_scene(16, 16, [(S.solid_tag(0), [S.rectangle_pairs(2.0, 2.0, 12.0, 12.0)])],
                       [[0.3, 0.6, 0.9, 0.5]], background=(1.0, 1.0, 1.0, 1.0), name="square")
    target = np.zeros((16, 16, 4), np.float32)
    pairs = sq.pairs.reshape(-1, 4).astype(np.float64)
    rgba = np.asarray(sq.substances, np.float64)
    h = 1e-4
    lp = O.diff_loss(sq, pairs, rgba + np.array([[h, 0, 0, 0]]), target, (5, 5, 6, 6))[0]
    lm = O.diff_loss(sq, pairs, rgba - np.array([[h, 0, 0, 0]]), target, (5, 5, 6, 6))[0]
    pix = 0.5 * 0.3 + 0.5 * 1.0
    assert abs((lp - lm) / (2 * h) - 2 * pix * 0.5) < 1e-8


def test_fd_samples_are_finite():
    sc, target = A.case(96, 60, 21)
    smp = A.fd_samples(sc, target, n=8, seed=1)
    assert len(smp) == 8 and all(np.isfinite(v) for *_, v in smp)


def test_fp64_adjoint_matches_finite_differences():
    """The adjoint's algorithm restated in FP64 (go_diff_grad: line weights,
    span-boundary terms, end points, knob splits, colour walk) against
    central differences of the FP64 forward on 64 sampled parameters."""
    sc, target = A.case(128, 40, 63)
    _, dp, dc = O.diff_grad(sc, sc.pairs, sc.substances, target)
    samples = A.fd_samples(sc, target, n=64, seed=3)
    assert len(samples) == 64
    got, fd, err, bad = A.compare(samples, dp, dc)
    assert not bad.any(), [(s, g) for s, g, b in zip(samples, got, bad) if b]


def test_forward_depends_on_column_segments():
    """Two lines of one shape cross inside the span y in [20, 20.26] of
    column 12.  A column that starts above the crossing keeps them in the
    order they became active (the reference's sweep does not reorder active
    records, Kernels.cl:1007-1124) and takes a section of negative width;
    a segment starting at y = 20 trims both at its top and sorts them by top
    x, the geometric order.  That pixel, and only it, differs; every other
    pixel is the same for 4-, 8- and 32-row segments."""
    sc, _ = A.case(32, 1, 33)
    target = np.zeros((32, 32, 4), np.float32)
    pairs = sc.pairs.reshape(-1, 4).astype(np.float64)
    rgba = np.asarray(sc.substances, np.float64)
    _, whole, _ = O.diff_loss(sc, pairs, rgba, target, (0, 0, 32, 32), want_image=True)
    for seg, differing in ((8, set()), (4, {(12, 20)})):
        img = np.zeros_like(whole)
        for y0 in range(0, 32, seg):
            img[y0:y0 + seg] = O.diff_loss(sc, pairs, rgba, target, (0, y0, 32, y0 + seg), want_image=True)[1]
        ys, xs = np.nonzero(np.abs(img - whole).max(axis=2) > 1e-12)
        assert set(zip(xs.tolist(), ys.tolist())) == differing, seg
        # the forward's own segmenting gives the same image in one call
        np.testing.assert_array_equal(O.diff_loss(sc, pairs, rgba, target, (0, 0, 32, 32), want_image=True,
                                                  seg=seg)[1], img)
    # a negative-width section in the whole column, none in the 4-row segment
    assert whole[20, 12, 0] < 0.95 < img[20, 12, 0]


def test_fp64_adjoint_on_short_segments():
    """With 4-row column segments (32-pixel tiles rendered by 256 threads)
    the adjoint's FP64 restatement is the gradient of the segmented forward:
    central differences of that forward agree on 32 sampled parameters,
    those of the one-segment forward do not."""
    sc, target = A.case(256, 150, 41)
    _, dp, dc = O.diff_grad(sc, sc.pairs, sc.substances, target, seg=4)
    samples = A.fd_samples(sc, target, n=32, seed=5, seg=4)
    assert len(samples) == 32
    got, fd, err, bad = A.compare(samples, dp, dc)
    assert not bad.any(), [(s, g) for s, g, b in zip(samples, got, bad) if b]
    whole = A.fd_samples(sc, target, n=32, seed=5)
    assert A.compare(whole, dp, dc)[3].any()


def test_segment_rows_of_the_tile_tree():
    import gudni_amd as G
    sc, _ = A.case(256, 150, 41)
    assert [A.segment_rows(sc, G.default_spec(max_tile_size=t)) for t in (256, 64, 32)] == [128, 16, 4]


def test_fp64_adjoint_closed_forms():
    """A translucent square over white: the colour gradient of one pixel is
    2 (pix - t) a per channel; moving the right edge by dx changes the loss by
    the edge column's colour difference times its height."""
    sq = S.build_scene(16, 16, [(S.solid_tag(0), [S.rectangle_pairs(2.0, 2.0, 12.5, 12.0)])],
                       [[0.3, 0.6, 0.9, 0.5]], background=(1.0, 1.0, 1.0, 1.0), name="square")
    target = np.zeros((16, 16, 4), np.float32)
    _, dp, dc = O.diff_grad(sq, sq.pairs, sq.substances, target)
    # d/dr = sum over pixels of 2 pix_r * a * coverage: 100 full pixels with
    # pix_r = 0.65 and the ten half-covered pixels of column 12, pix_r = 0.825
    r = float(np.float32(0.3))   # the substance table is FP32
    expect_r = 100 * 2 * (0.5 * r + 0.5) * 0.5 + 10 * 2 * (0.25 * r + 0.75) * 0.5 * 0.5
    assert abs(dc[0, 0] - expect_r) < 1e-9
    # the right edge x = 12.5 of rows 2..12: half-covered column 12
    col = np.float32([0.3, 0.6, 0.9]).astype(np.float64)
    inside = np.append(0.5 * col + 0.5, 1.0)
    outside = np.ones(4)
    g_col12 = 2 * (0.5 * inside + 0.5 * outside)          # dL/dpixel in column 12
    expect = float(g_col12 @ (inside - outside)) * 10.0     # ten rows
    pts = sq.pairs.reshape(-1, 4)
    right = [i for i in range(len(pts)) if abs(pts[i, 0] - 12.5) < 1e-6]
    got = sum(dp[i, 0] for i in right) + sum(dp[i, 2] for i in range(len(pts)) if abs(pts[i, 2] - 12.5) < 1e-6)
    assert abs(got - expect) < 1e-9, (got, expect)


def fp64_image_ex(scene, seg_map=None, filt=0):
    target = np.zeros((scene.height, scene.width, 4), np.float32)
    _, img, _ = O.diff_loss_ex(scene, scene.pairs.reshape(-1, 4), np.asarray(scene.substances, np.float64),
                               target, (0, 0, scene.width, scene.height), seg_map=seg_map, filter=filt,
                               want_image=True)
    return img


def test_segment_map_of_one_height_equals_uniform_segments():
    sc = polygon_scene(60, 128, 9)
    target = np.zeros((sc.height, sc.width, 4), np.float32)
    pairs, rgba = sc.pairs.reshape(-1, 4), np.asarray(sc.substances, np.float64)
    a = O.diff_loss(sc, pairs, rgba, target, (0, 0, 128, 128), want_image=True, seg=16)
    b = O.diff_loss_ex(sc, pairs, rgba, target, (0, 0, 128, 128), want_image=True,
                       seg_map=np.full((128, 128), 16, np.int32))
    assert a[0] == b[0] and a[2] == b[2]
    np.testing.assert_array_equal(a[1], b[1])


def test_segment_map_follows_a_mixed_tile_tree():
    """A tile tree of mixed tile sizes cuts columns into segments of
    different heights; the FP64 forward cut by the tree's segment map gives
    the oracle's pixels, which are rendered through those very tiles."""
    import gudni_amd as G
    rng = np.random.default_rng(17)
    shapes, subs = [], []
    for i in range(700):   # crowded in one corner, sparse elsewhere
        k = int(rng.integers(3, 9))
        crowded = i < 600
        c = rng.uniform(0, 120 if crowded else 512, 2)
        r = rng.uniform(2, 10) if crowded else rng.uniform(5, 40)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        shapes.append((S.solid_tag(i), [S.straight_pairs([(float(c[0] + r * np.cos(a)), float(c[1] + r * np.sin(a)))
                                                           for a in ang])]))
        subs.append([float(v) for v in rng.uniform(0, 1, 4)])
    sc = S.build_scene(512, 512, shapes, subs, name="crowded corner")
    spec = G.default_spec()
    m = A.segment_map(sc, spec)
    assert len(np.unique(m)) >= 2, "the tree should mix segment heights"
    ref = H.oracle_render(sc, spec, want_f64=True)
    d = np.abs(fp64_image_ex(sc, m) - ref["f64"].reshape(sc.height, sc.width, 4)).max(axis=2)
    assert np.median(d) < 1e-7
    assert (d > 1e-5).mean() < 1e-3, f"{(d > 1e-5).sum()} pixels differ by more than 1e-5"


def test_pictures_match_oracle():
    """Picture substances sampled nearest (the reference's getPicturePixel)
    and bilinear, at the pixel's integer position, as the oracle samples
    them."""
    import gudni_amd as G
    sc = S.batch_scene(192, 150, 23, gradient_share=0, picture_share=4, texture_side=64)
    kinds = np.asarray(sc.tags, np.uint64) & np.uint64(0xC000000000000000)
    assert (kinds == np.uint64(S.TAG_PICTURE)).sum() >= 20
    for filt in (0, 1):
        ref = H.oracle_render(sc, G.default_spec(texture_filter=filt), want_f64=True)
        d = np.abs(fp64_image_ex(sc, None, filt) - ref["f64"].reshape(sc.height, sc.width, 4)).max(axis=2)
        assert np.median(d) < 1e-7, filt
        assert (d > 1e-5).mean() < 2e-2 and d.max() < 2e-2, filt
    # the two filters give different images
    a, b = fp64_image_ex(sc, None, 0), fp64_image_ex(sc, None, 1)
    assert np.abs(a - b).max() > 0.05


def test_fd_threads_give_the_same_samples():
    sc, target = A.case(96, 60, 21)
    a = A.fd_samples(sc, target, n=10, seed=2, threads=1)
    b = A.fd_samples(sc, target, n=10, seed=2, threads=8)
    assert a == b and len(a) == 10


def compound_mix_scene(seed=4, size=128, n=60):
    """Shapes sharing substances, subtract compounds (cSubtract, Figure/
    ShapeTree.hs:84-166) and solid colours, on straight edges."""
    rng = np.random.default_rng(seed)
    shapes, subs = [], []
    for i in range(n):
        k = int(rng.integers(3, 8))
        c = rng.uniform(0, size, 2)
        r = rng.uniform(6, 30)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        pts = [(float(c[0] + r * np.cos(a)), float(c[1] + r * np.sin(a))) for a in ang]
        sid = int(rng.integers(0, n // 3))          # about three shapes per substance
        subtract = rng.uniform() < 0.3
        shapes.append((S.solid_tag(sid, subtract=subtract), [S.straight_pairs(pts)]))
    for _ in range(n // 3):
        subs.append([float(v) for v in rng.uniform(0, 1, 4)])
    return S.build_scene(size, size, shapes, subs, name="compound mix")


def test_fp64_forward_merges_compounds_like_the_oracle():
    """determineColor's merge of shapes with one substance id and its
    subtract compounds (Kernels.cl:1447-1513) in the FP64 forward gives the
    oracle's pixels."""
    from tests.test_gpu_parity import compound_scene
    for sc in (compound_scene(), compound_mix_scene(), compound_mix_scene(9, 96, 45)):
        ref = H.oracle_render(sc, want_f64=True)
        d = np.abs(fp64_image(sc) - ref["f64"].reshape(sc.height, sc.width, 4)).max(axis=2)
        assert np.median(d) < 1e-7, sc.name
        assert (d > 1e-5).mean() < 1e-3, (sc.name, (d > 1e-5).sum())


def test_fp64_adjoint_with_compounds_matches_finite_differences():
    """The FP64 adjoint with determineColor's merge walk (shapes sharing a
    substance, subtract compounds) against central differences; a colour
    sample's window covers every shape of its substance."""
    import copy
    base = copy.deepcopy(S.batch_scene(160, 90, 47, gradient_share=0))
    n_subs = 25
    base.tags = np.array([S.solid_tag(i % n_subs, subtract=(i % 4 == 1)) for i in range(base.n_shapes)],
                         np.uint64)
    base.substances = np.ascontiguousarray(np.asarray(base.substances, np.float32)[:n_subs])
    target = H.oracle_render(base, want_f64=False)["f32"].reshape(160, 160, 4).astype(np.float32)
    sc = A.perturbed(base, 1047)
    _, dp, dc = O.diff_grad(sc, sc.pairs, sc.substances, target, seg=256)
    samples = A.fd_samples(sc, target, n=64, seed=5, seg=256)
    assert len(samples) == 64 and sum(k == "rgba" for k, *_ in samples) >= 8
    got, fd, err, bad = A.compare(samples, dp, dc)
    assert not bad.any(), [(s, g) for s, g, b in zip(samples, got, bad) if b]
