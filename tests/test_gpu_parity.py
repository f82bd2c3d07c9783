"""Parity of the HIP path (through the C ABI) with the oracle.

Integer/byte outputs (records, headers, their order, shape state, BGRA
words) must be bit-identical; pre-quantisation colours must lie within 1e-5
of the oracle's FP64 integration of the same sections.
"""
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

COVERAGE_TOL = 1e-5


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


def comb_scene(teeth=100, size=256):
    """One zig-zag polygon whose middle columns cross 2*teeth edges inside a
    single 256-pixel column segment: the live queue outgrows the fast pass's
    rings and the unit is re-run by the full-capacity pass."""
    pts = [(10.0, 10.0)]
    for i in range(teeth):
        pts.append((200.5, 10.0 + 2.2 * i + 0.55))
        pts.append((20.5, 10.0 + 2.2 * i + 1.1))
    pts.append((10.0, 10.0 + 2.2 * teeth + 1.0))
    return S.build_scene(size, size, [(S.solid_tag(0), [S.straight_pairs(pts)])],
                         [[0.2, 0.4, 0.8, 0.9]], name="comb")


def textured_scene():
    rng = S.SplitMix64(11)
    tex = (rng.next(32 * 32) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.uint8)
    shapes = [(S.solid_tag(0), [S.rectangle_pairs(3.3, 4.1, 40.7, 30.2)]),
              (S.picture_tag(1), [S.straight_pairs([(10.2, 2.5), (60.1, 20.3), (30.4, 55.9)])]),
              (S.picture_tag(2), [S.rectangle_pairs(35.5, 35.5, 63.0, 63.5)])]
    subs = [[0.9, 0.1, 0.1, 0.6], S.picture_substance(0), S.picture_substance(1)]
    uses = [(-8.0, -3.0, 32, 32, 0, 1.0), (-30.0, -30.0, 32, 32, 0, 0.5)]
    return S.build_scene(64, 64, shapes, subs, tex, uses, background=(0.1, 0.2, 0.3, 1.0),
                         name="textured")


def compound_scene():
    """openRectangle-style compound (Layout/Draw.hs:96-103): a rectangle with a
    subtracted inner rectangle of the same substance, over a second shape."""
    shapes = [(S.solid_tag(0), [S.rectangle_pairs(4.25, 4.25, 28.75, 28.75)]),
              (S.solid_tag(0, subtract=True), [S.rectangle_pairs(8.5, 8.5, 24.5, 24.5)]),
              (S.solid_tag(1), [S.straight_pairs([(0.5, 16.5), (31.5, 2.5), (31.5, 30.5)])])]
    return S.build_scene(32, 32, shapes, [[0.0, 0.5, 1.0, 0.8], [1.0, 0.8, 0.0, 0.5]],
                         name="compound")


def check_frames(scene, gpu_u32, gpu_f32, ref):
    bad = H.first_mismatch(gpu_u32, ref["out"])
    assert bad is None, f"{scene.name}: first BGRA mismatch at {bad}"
    err = np.abs(gpu_f32.astype(np.float64) - ref["f64"]).max() if gpu_f32 is not None else 0.0
    assert err <= COVERAGE_TOL, f"{scene.name}: coverage error {err} > {COVERAGE_TOL}"
    return err


@pytest.mark.parametrize("name", ["tinySquare", "compound", "textured", "uniform_small"])
def test_small_scenes_level2(rasterizer, name):
    scene = {"tinySquare": H.tiny_square, "compound": compound_scene, "textured": textured_scene,
             "uniform_small": lambda: S.uniform_512(200, 128, 3)}[name]()
    say(f"level2 {name} {scene.width}x{scene.height}")
    ref = H.oracle_render(scene)
    assert ref["rc"] == 0
    u32, f32 = H.gpu_render_level2(rasterizer, scene)
    check_frames(scene, u32, f32, ref)


@pytest.mark.parametrize("mode", [0, 1])
def test_stochastic_field(mode):
    """The reference's stochastic anti-aliasing (initRandomField / getRandom,
    Kernels.cl:1597-1613, 1738-1741), inert at the default factor 0: with a
    non-zero factor every section's area is scaled by 1 + r * factor, r walking
    the random field from a per-column start.  Frames equal the oracle's in
    both coverage modes (and differ from the factor-0 frames)."""
    spec = G.default_spec(stochastic_factor=0.35, coverage_mode=mode)
    r = G.Rasterizer(spec)
    try:
        for scene in (S.uniform_512(200, 128, 3), textured_scene(), gradient_scene()):
            say(f"stochastic mode {mode}: {scene.name}")
            ref = H.oracle_render(scene, spec)
            assert ref["rc"] == 0
            plain = H.oracle_render(scene, G.default_spec(coverage_mode=mode), want_f64=False)
            assert (plain["out"] != ref["out"]).any()
            u32, f32 = H.gpu_render_level2(r, scene, spec)
            check_frames(scene, u32, f32, ref)
    finally:
        r.close()


def test_tiny_square_values(rasterizer):
    """Closed form: pixel (0,0) is 0.81 red over white, (1,1) pure red."""
    scene = H.tiny_square()
    say("tinySquare closed form")
    u32, f32 = H.gpu_render_level2(rasterizer, scene)
    assert u32[1, 1] == 0xFFFF0000
    assert u32[5, 5] == 0xFFFFFFFF
    assert abs(f32[0, 0, 1] - (1.0 - 0.81)) < 1e-5


def dump_parity(rasterizer, scene, job_ids, cap=64, spec=None):
    spec, geo, jobs = G.prepare(scene, spec)
    rasterizer.upload_common(geo.geo, scene)
    ospec = H.oracle_spec(spec)
    og = O.Geometry(scene)
    oj = O.JobList(ospec, og, scene.width, scene.height)
    for j in job_ids(jobs.n_jobs):
        say(f"{scene.name} dump job {j}")
        tiles, shapes = jobs.job(j)
        d = rasterizer.dump_job(shapes, tiles, scene.width, scene.height, jobs.n_jobs - 1 - j, cap=cap)
        rc, _, od = O.raster_job_dump(scene, ospec, og, oj, j, cap=cap)
        assert rc == 0
        np.testing.assert_array_equal(d["qlen"], od["qlen"])
        assert d["qlen"].max() <= cap
        for k in ("pre_thr", "thr"):
            np.testing.assert_array_equal(d[k].view(np.uint32), od[k].view(np.uint32), err_msg=k)
        for k in ("pre_hdr", "hdr", "shape_bits", "stack"):
            np.testing.assert_array_equal(d[k], od[k], err_msg=k)
        bits = od["shape_bits"].astype(np.int64)
        mask = np.arange(spec.max_shapes)[None, :] < np.minimum(bits, spec.max_shapes)[:, None]
        np.testing.assert_array_equal(np.where(mask, d["shape_indices"], 0),
                                      np.where(mask, od["shape_indices"], 0))


def test_uniform_512_dump_parity(rasterizer):
    """Per-column records after generate and after sort, shape bits, initial
    stack and shape indices: bit-identical to the oracle on two jobs.  the 512^2 uniform scene's
    tiles sit at the 8-pixel floor with ~750 shapes each (general path)."""
    dump_parity(rasterizer, S.uniform_512(), lambda n: (0, n - 1))


def test_blobs_1024_dump_parity(rasterizer):
    """The same on a 1024^2 version of the 4096^2 blob scene, whose tiles hold <= 126 shapes
    (compact path: records labelled by tile-local shape index internally and
    relabelled to the reference's shape bits here)."""
    dump_parity(rasterizer, S.blobs_scaled(1024, 12500, 7), lambda n: range(n), cap=96)


@pytest.mark.parametrize("bits", [0, 2, 5])
def test_sort_radix_dump_parity(bits):
    """The radix pass ahead of the per-column insertion sort leaves the sorted
    records bit-identical to the oracle's for every digit width: 256-row
    columns of a sparse scene (tens of records each, compact path) and the
    8-pixel-floor tiles of the 512^2 uniform scene (general path)."""
    spec = G.default_spec(sort_radix_bits=bits)
    r = G.Rasterizer(spec)
    try:
        dump_parity(r, S.batch_scene(512, 700, 21, gradient_share=0), lambda n: range(n), cap=256,
                    spec=spec)
        dump_parity(r, S.uniform_512(), lambda n: (n - 1,), spec=spec)
    finally:
        r.close()


def test_uniform_512_level1_and_level2(rasterizer):
    scene = S.uniform_512()
    say("uniform 512 oracle")
    ref = H.oracle_render(scene)
    assert ref["rc"] == 0
    say("uniform 512 level1 (per job, host buffers)")
    u1 = H.gpu_render_level1(rasterizer, scene)
    assert H.first_mismatch(u1, ref["out"]) is None
    say("uniform 512 level2 (whole frame, device buffers)")
    u2, f2 = H.gpu_render_level2(rasterizer, scene)
    err = check_frames(scene, u2, f2, ref)
    say(f"uniform 512 coverage max |f32 - f64| = {err:.3g}")


def test_blobs_1024(rasterizer):
    scene = S.blobs_scaled(1024, 12500, 7)
    say("blobs 1024^2")
    ref = H.oracle_render(scene, nthreads=16)
    assert ref["rc"] == 0
    u2, f2 = H.gpu_render_level2(rasterizer, scene)
    err = check_frames(scene, u2, f2, ref)
    say(f"blobs-1024 coverage max err {err:.3g}")


@pytest.mark.parametrize("tile,filt,mode", [(256, 0, 0), (32, 1, 0), (256, 0, 1)])
def test_blobs_4096_full_frame(rasterizer, tile, filt, mode):
    """The full 4096^2, 200k-shape workload: every BGRA word equals the
    oracle's, with the pinned spec, with the benchmark's (32-pixel tiles,
    bilinear texture sampling), and in the reference's own arithmetic
    (coverage mode REFERENCE)."""
    scene = S.blobs_4096()
    spec = G.default_spec(max_tile_size=tile, texture_filter=filt, coverage_mode=mode)
    say(f"blobs 4096 oracle (16 threads), tile {tile} filter {filt} mode {mode}")
    ref = H.oracle_render(scene, spec, want_f64=False, nthreads=16)
    assert ref["rc"] == 0
    say("blobs 4096 gpu")
    r = rasterizer if tile == 256 and filt == 0 and mode == 0 else G.Rasterizer(spec)
    try:
        u2, _ = H.gpu_render_level2(r, scene, spec)
    finally:
        if r is not rasterizer:
            r.close()
    bad = H.first_mismatch(u2, ref["out"])
    assert bad is None, f"blobs 4096 first mismatch {bad}"


def test_retry_pass_comb(rasterizer):
    """Queues longer than the fast ring go through the full-capacity pass and
    still match the oracle bit for bit."""
    scene = comb_scene()
    say("comb (retry pass)")
    ref = H.oracle_render(scene)
    assert ref["rc"] == 0
    u2, f2 = H.gpu_render_level2(rasterizer, scene)
    check_frames(scene, u2, f2, ref)


def test_queue_overflow_is_an_error():
    """maxThresholds smaller than a column's queue: the reference would
    overwrite the neighbouring column (Kernels.cl:375-380); both the oracle and
    the HIP path report it."""
    scene = comb_scene(teeth=40)
    spec = G.default_spec(max_thresholds=16)
    say("overflow: oracle")
    ref = H.oracle_render(scene, spec)
    assert ref["rc"] == -1
    say("overflow: gpu")
    r = G.Rasterizer(spec)
    try:
        with pytest.raises(G.GudniError) as ei:
            H.gpu_render_level2(r, scene, spec)
        assert ei.value.code == -3
    finally:
        r.close()


@pytest.mark.parametrize("filt", [0, 1])   # nearest, bilinear
def test_texel_bytes_are_exact(filt):
    """Every byte value of a picture becomes byte / 255 correctly rounded (the
    kernel multiplies by the reciprocal and corrects with one FMA; the oracle
    divides): a 16x16 texture holding all 256 values in each channel (alpha
    included), sampled by a canvas-filling rectangle over an opaque
    background, gives pre-quantisation colours equal to the oracle's bit for
    bit."""
    i = np.arange(256, dtype=np.uint32)
    tex = np.stack([i, 255 - i, (i * 7) & 255, (i * 13 + 5) & 255], axis=1).astype(np.uint8).reshape(-1)
    shapes = [(S.picture_tag(0), [S.rectangle_pairs(0.0, 0.0, 24.0, 24.0)])]
    scene = S.build_scene(24, 24, shapes, [S.picture_substance(0)], tex, [(0.0, 0.0, 16, 16, 0, 1.0)],
                          background=(0.25, 0.5, 0.75, 1.0), name="texel bytes")
    spec = G.default_spec(texture_filter=filt)
    ref = H.oracle_render(scene, spec)
    assert ref["rc"] == 0
    r = G.Rasterizer(spec)
    try:
        u32, f32 = H.gpu_render_level2(r, scene, spec)
    finally:
        r.close()
    want = np.asarray(ref["f32"], np.float32).reshape(f32.shape)
    assert np.array_equal(f32.view(np.uint32), want.view(np.uint32)), \
        np.argwhere(f32.view(np.uint32) != want.view(np.uint32))[:5]
    assert H.first_mismatch(u32, ref["out"]) is None


def test_level1_rejects_malformed_references(rasterizer):
    """A job's shape references are checked on the host before anything runs
    (in parallel chunks for a large job): a substance id outside the table or
    a strand outside the geometry heap is an error naming the first malformed
    reference, and the same job with valid references renders."""
    scene = S.uniform_512(300, 256, 5)
    _, geo, jobs = G.prepare(scene)
    rasterizer.upload_common(geo.geo, scene)
    tiles, shapes = jobs.job(0)
    ref_t = np.dtype([("tag", "<u8"), ("geo", "<u4"), ("n", "<u4")])
    refs = np.frombuffer(np.asarray(shapes, np.uint8).tobytes(), ref_t)
    big = np.resize(refs, 200_000)   # the tiles read the first len(refs) of them
    heap_units = np.asarray(geo.geo, np.uint8).size // 16

    def job_with(i_sid, i_geo):
        bad = big.copy()
        if i_sid is not None:
            bad["tag"][i_sid] = (int(bad["tag"][i_sid]) & ~int(S.TAG_ID_MASK)) | (len(scene.substances) + 7)
        if i_geo is not None:
            bad["geo"][i_geo] = heap_units + 3
            bad["n"][i_geo] = 1
        return np.frombuffer(bad.tobytes(), np.uint8)

    out = np.zeros(scene.width * scene.height, np.uint32)
    for i_sid, i_geo, want in [(150_000, 190_000, "substance id"), (180_000, 120_000, "geometry heap"),
                               (70_000, None, "substance id"), (None, 199_999, "geometry heap")]:
        say(f"malformed references at {i_sid}, {i_geo}")
        with pytest.raises(G.GudniError) as ei:
            rasterizer.raster_job(job_with(i_sid, i_geo), tiles, scene.background, scene.width, scene.height,
                                  0, jobs.n_jobs - 1, out)
        assert ei.value.code == -1 and want in str(ei.value), str(ei.value)
    good = rasterizer.raster_job(job_with(None, None), tiles, scene.background, scene.width, scene.height,
                                 0, jobs.n_jobs - 1, out)
    alone = rasterizer.raster_job(np.asarray(shapes, np.uint8), tiles, scene.background, scene.width,
                                  scene.height, 0, jobs.n_jobs - 1, np.zeros_like(out))
    assert np.array_equal(good, alone)


def test_empty_and_offcanvas(rasterizer):
    """A scene whose only shape lies off the canvas is culled (excludeBox):
    every pixel is the background."""
    scene = S.build_scene(40, 24, [(S.solid_tag(0), [S.rectangle_pairs(50, 50, 60, 60)])],
                          [[1, 0, 0, 1]], background=(0.25, 0.5, 0.75, 1.0), name="offcanvas")
    say("offcanvas")
    ref = H.oracle_render(scene)
    u2, f2 = H.gpu_render_level2(rasterizer, scene)
    check_frames(scene, u2, f2, ref)
    assert np.all(u2 == u2[0, 0])


def test_nonsquare_canvas(rasterizer):
    scene = S.uniform_512(300, 200, 13)
    scene.height = 120
    say("non-square 200x120")
    ref = H.oracle_render(scene)
    u2, f2 = H.gpu_render_level2(rasterizer, scene)
    check_frames(scene, u2, f2, ref)


def gradient_scene():
    """Overlapping translucent gradient, picture and solid shapes."""
    rng = S.SplitMix64(23)
    tex = (rng.next(24 * 24) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.uint8)
    cols = [[1.0, 0.2, 0.1, 0.9], [0.1, 0.8, 0.3, 0.5], [0.2, 0.3, 1.0, 0.7], [0.9, 0.9, 0.1, 1.0]]
    grads = [S.gradient_record(0, (5.0, 6.0), (70.0, 40.0), [0.0, 0.4, 0.7, 1.0], cols),
             S.gradient_record(1, (40.0, 45.0), (30.0, 0.0), [0.0, 1.0], cols[:2])]
    shapes = [(S.gradient_tag(0), [S.straight_pairs([(3.5, 4.2), (75.1, 12.7), (40.3, 70.9)])]),
              (S.gradient_tag(1), [S.rectangle_pairs(20.25, 22.5, 66.75, 71.0)]),
              (S.picture_tag(2), [S.straight_pairs([(10.0, 50.0), (50.5, 30.5), (44.0, 78.0)])]),
              (S.solid_tag(3), [S.rectangle_pairs(30.5, 5.5, 60.5, 35.5)])]
    subs = [S.gradient_substance(0), S.gradient_substance(1), S.picture_substance(0),
            [0.3, 0.6, 0.9, 0.6]]
    return S.build_scene(80, 80, shapes, subs, tex, [(-7.5, -13.25, 24, 24, 0, 0.8)],
                         background=(0.95, 0.9, 0.8, 1.0), name="gradients", gradients=grads)


@pytest.mark.parametrize("filt", [0, 1])
def test_gradients_and_texture_filters(filt):
    """Gradient substances and nearest / bilinear pictures: BGRA bit-identical
    to the oracle, colours within the coverage tolerance."""
    spec = G.default_spec(texture_filter=filt)
    r = G.Rasterizer(spec)
    try:
        for scene in (gradient_scene(), S.blobs_scaled(512, 4000, 17)):
            say(f"filter {filt}: {scene.name}")
            ref = H.oracle_render(scene, spec)
            assert ref["rc"] == 0
            u2, f2 = H.gpu_render_level2(r, scene, spec)
            check_frames(scene, u2, f2, ref)
    finally:
        r.close()


def test_rgba8_output_format():
    """GUDNI_FORMAT_RGBA8 writes the reference's BGRA words' channel bytes in
    R,G,B,A order, on the whole-frame, per-job and batch paths (the batch's
    captured graph follows a format change); BGRA stays the oracle's."""
    import torch
    scene = textured_scene()
    ref = H.oracle_render(scene, want_f64=False)
    assert ref["rc"] == 0
    bgra = np.asarray(ref["out"], np.uint32)
    want = (bgra & 0xFF00FF00) | ((bgra >> 16) & 0xFF) | ((bgra & 0xFF) << 16)
    r = G.Rasterizer()
    try:
        say("rgba8: whole frame")
        r.set_output_format(G.FORMAT_RGBA8)
        u2, _ = H.gpu_render_level2(r, scene)
        assert H.first_mismatch(u2, want) is None
        say("rgba8: per job")
        u1 = H.gpu_render_level1(r, scene)
        assert H.first_mismatch(u1, want) is None
        say("rgba8: batch with graph, format switched between replays")
        r.set_stream(torch.cuda.current_stream().cuda_stream)
        b = G.Batch.from_outlines(r, [scene, H.tiny_square()])
        for fmt, expect in ((G.FORMAT_RGBA8, want), (G.FORMAT_BGRA8, bgra), (G.FORMAT_RGBA8, want)):
            r.set_output_format(fmt)
            b.render(use_graph=True)
            b.check()
            assert H.first_mismatch(b.frame(0), expect) is None, fmt
        b.close()
        with pytest.raises(G.GudniError):
            r.set_output_format(7)
    finally:
        r.close()


def test_level1_job_touches_only_its_tiles(rasterizer):
    """gudni_raster_job writes a job's pixels into the host frame and leaves
    every other pixel as it was (it moves only the window of the job's
    tiles): one job of the 512^2 uniform scene into a frame of sentinels."""
    scene = S.uniform_512()
    _, geo, jobs = G.prepare(scene)
    rasterizer.upload_common(geo.geo, scene)
    ref = H.oracle_render(scene, want_f64=False)["out"]
    sentinel = np.uint32(0xDEADBEEF)
    for j in (0, jobs.n_jobs // 2, jobs.n_jobs - 1):
        tiles, shapes = jobs.job(j)
        say(f"level1 single job {j} of {jobs.n_jobs}")
        out = np.full(scene.width * scene.height, sentinel, np.uint32)
        rasterizer.raster_job(shapes, tiles, scene.background, scene.width, scene.height, 0,
                              jobs.n_jobs - 1 - j, out)
        frame = out.reshape(scene.height, scene.width)
        mine = np.zeros(frame.shape, bool)
        for b in np.frombuffer(tiles.tobytes(), np.int32).reshape(-1, 8)[:, :4]:
            mine[max(0, b[1]):min(scene.height, b[3]), max(0, b[0]):min(scene.width, b[2])] = True
        assert mine.any()
        assert np.array_equal(frame[mine], ref[mine]), f"job {j}: its pixels differ from the oracle"
        assert np.all(frame[~mine] == sentinel), f"job {j}: pixels outside its tiles changed"


def test_level1_job_writes_below_a_short_tile_box(rasterizer):
    """A column writes 2^dd rows from its origin, clipped to the canvas but
    not to the tile box (initTileState, Kernels.cl:1692-1722).  A 256-pixel
    tile whose box is cut to 100 rows still writes all 256 rows of each
    column, and gudni_raster_job returns every one of them to the host
    frame, as the oracle writes them."""
    scene = S.build_scene(256, 256, [(S.solid_tag(0), [S.rectangle_pairs(10.5, 20.25, 200.5, 240.75)])],
                          [[0.2, 0.6, 1.0, 1.0]], background=(0.9, 0.8, 0.7, 1.0), name="short box")
    _, geo, jobs = G.prepare(scene)
    assert jobs.n_jobs == 1
    tiles, shapes = jobs.job(0)
    t = np.frombuffer(tiles.tobytes(), H.TILE_DT).copy()
    assert t.size == 1 and tuple(t["box"][0]) == (0, 0, 256, 256) and t["hd"][0] == 8 and t["vd"][0] == 8
    t["box"][0][3] = 100
    short = t.view(np.uint8)
    say("level1 job with a tile box cut to 100 of its 256 rows")
    rasterizer.upload_common(geo.geo, scene)
    out = np.full(scene.width * scene.height, np.uint32(0xDEADBEEF), np.uint32)
    rasterizer.raster_job(shapes, short, scene.background, scene.width, scene.height, 0, 0, out)
    ospec = H.oracle_spec(G.default_spec())
    ogeo = O.Geometry(scene)
    assert np.array_equal(np.frombuffer(ogeo.geo.tobytes(), np.uint8), np.frombuffer(geo.geo.tobytes(), np.uint8))

    class OneJob:
        n_jobs = 1

        def job(self, i):
            return short, np.ascontiguousarray(shapes, np.uint8)

    rc, ref, _ = O.raster_job_dump(scene, ospec, ogeo, OneJob(), 0, cap=1)
    assert rc == 0
    frame = out.reshape(scene.height, scene.width)
    assert np.array_equal(frame, ref), H.first_mismatch(frame, ref)
    assert not np.any(frame[100:] == np.uint32(0xDEADBEEF)), "rows below the box did not reach the host"


def test_level1_job_invalidates_the_uploaded_frame():
    """A Level-1 job writes its records into the buffers an uploaded frame
    renders from, so a frame render after it must refuse (until the frame is
    uploaded again) rather than render the job's records; after a fresh
    upload the frame equals the oracle's again."""
    import torch
    r = G.Rasterizer()
    try:
        scene = S.uniform_512(300, 256, 5)
        _, geo, jobs = G.prepare(scene)
        r.upload_common(geo.geo, scene)
        r.frame_upload(jobs, scene)
        tiles, shapes = jobs.job(0)
        out = np.zeros(scene.width * scene.height, np.uint32)
        say("frame_upload, then a level-1 job, then frame_render")
        r.raster_job(shapes, tiles, scene.background, scene.width, scene.height, 0, jobs.n_jobs - 1, out)
        dev = torch.zeros(scene.width * scene.height, dtype=torch.int32, device="cuda")
        with pytest.raises(G.GudniError) as ei:
            r.frame_render(dev.data_ptr())
        assert ei.value.code == -1
        got, _ = H.gpu_render_level2(r, scene)
        ref = H.oracle_render(scene, want_f64=False)["out"]
        assert H.first_mismatch(got, ref) is None
    finally:
        r.close()
