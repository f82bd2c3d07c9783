"""The adjoint through the C ABI: its forward image equals the oracle's
render of the same scene within the coverage tolerance, its loss equals the
FP64 forward's, and its gradients match FP64 central finite differences on
64 sampled parameters (points h = 1e-2 px, colours h = 1e-3), samples where a
discrete decision flips rejected.  Several scenes in one launch give each
scene's own gradients."""
import sys

import numpy as np
import pytest

import gudni_amd as G
from oracle import oracle as O
from tests import adjoint_cases as A
from tests import helpers as H

pytestmark = pytest.mark.gpu

COVERAGE_TOL = 1e-5


def say(msg):
    print(f"[gpu-adjoint] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def adjoint_small():
    sc, target = A.case(256, 150, 41)
    return sc, target


def gpu_adjoint(r, sc, target):
    adj = G.Adjoint(r, [(sc, target)])
    try:
        adj.run()
        return adj.result(0, image=True)
    finally:
        adj.close()


def tiled(rasterizer, tile):
    """The shared context for the pinned tile size, else a context whose
    device tile tree uses `tile` (the autotuner's smaller sizes)."""
    if tile == 256:
        return rasterizer, None
    r = G.Rasterizer(G.default_spec(max_tile_size=tile))
    return r, r


@pytest.mark.parametrize("tile", [256, 64, 32, 16])
def test_adjoint_forward_matches_oracle(rasterizer, adjoint_small, tile):
    sc, target = adjoint_small
    say(f"adjoint forward vs oracle, tile {tile}")
    r, own = tiled(rasterizer, tile)
    try:
        loss, dp, dc, img = gpu_adjoint(r, sc, target)
    finally:
        if own:
            own.close()
    ref = H.oracle_render(sc, G.default_spec(max_tile_size=tile), want_f64=True)
    err = np.abs(img.astype(np.float64) - ref["f64"].reshape(img.shape)).max()
    assert err <= COVERAGE_TOL, f"forward image differs from the oracle by {err}"
    l64, _, _ = O.diff_loss(sc, sc.pairs.reshape(-1, 4), np.asarray(sc.substances, np.float64), target,
                            (0, 0, sc.width, sc.height))
    ref_loss = float(((ref["f64"].reshape(img.shape) - target) ** 2).sum())
    say(f"loss gpu {loss:.6f} oracle {ref_loss:.6f} fp64 forward {l64:.6f}")
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    assert np.all(np.isfinite(dp)) and np.all(np.isfinite(dc))


def fan_scene(k=50):
    """One polygon zig-zagging 2k times between x = 10.2 and 10.8 inside
    column 10 of a 64^2 canvas, each edge its own strand, the edges' tops
    staggered down the first 16 rows and their bottoms at row 63.5.  At 64-px
    tiles the column's top 16-row segment generates one record per edge,
    within the fast pass's 128-record rings, but its sweep splits every
    started record at each later top, so the live queue grows to about twice
    that before the split-off tops are popped."""
    from gudni_amd import scenes as S
    pts = []
    for i in range(k):
        pts.append((10.2, 0.5 + 0.3 * i))
        pts.append((10.8, 63.5))
    return S.build_scene(64, 64, [(S.solid_tag(0), [S.straight_pairs(pts)])], [[0.3, 0.6, 0.2, 0.8]],
                         name="fan")


def test_forward_sweep_outgrowing_its_ring_is_retried():
    """A unit whose records fit the fast pass's ring but whose forward sweep
    outgrows it (the sweep runs on a copy of the sorted records) goes to the
    full-capacity pass like one whose generation overflows: its pixels come
    out as the oracle's."""
    sc = fan_scene()
    target = np.zeros((sc.height, sc.width, 4), np.float32)
    say("adjoint forward of a sweep that outgrows its ring")
    r = G.Rasterizer(G.default_spec(max_tile_size=64))
    try:
        loss, dp, dc, img = gpu_adjoint(r, sc, target)
    finally:
        r.close()
    ref = H.oracle_render(sc, G.default_spec(max_tile_size=64), want_f64=True)
    err = np.abs(img.astype(np.float64) - ref["f64"].reshape(img.shape)).max()
    assert err <= COVERAGE_TOL, f"forward image differs from the oracle by {err}"
    assert np.all(np.isfinite(dp)) and np.all(np.isfinite(dc))


@pytest.mark.parametrize("tile", [256, 64, 32, 16])
def test_adjoint_gradients_match_finite_differences(rasterizer, adjoint_small, tile):
    sc, target = adjoint_small
    say(f"adjoint gradients, tile {tile}")
    r, own = tiled(rasterizer, tile)
    try:
        loss, dp, dc, _ = gpu_adjoint(r, sc, target)
    finally:
        if own:
            own.close()
    seg = A.segment_rows(sc, G.default_spec(max_tile_size=tile))
    say(f"finite differences (64 samples, FP64, {seg}-row column segments)")
    samples = A.fd_samples(sc, target, n=64, seed=5, seg=seg)
    assert len(samples) == 64
    got, fd, err, bad = A.compare(samples, dp, dc)
    for (kind, i, k, v), g, e, b in zip(samples, got, err, bad):
        say(f"{kind:5s} {i:5d}.{k} fd {v:+.6e} adjoint {g:+.6e} err {e:.2e}{' BAD' if b else ''}")
    assert bad.sum() == 0, f"{bad.sum()} of 64 gradients off"
    # and the same algorithm in FP64 over every parameter
    _, dp64, dc64 = O.diff_grad(sc, sc.pairs, sc.substances, target, seg=seg)
    scale_p = np.sqrt(np.mean(dp64 ** 2))
    assert np.abs(dp - dp64).max() <= 2e-2 * scale_p + 0.0, np.abs(dp - dp64).max() / scale_p
    np.testing.assert_allclose(dc, dc64, rtol=1e-3, atol=1e-4 * np.sqrt(np.mean(dc64 ** 2)))


def test_batched_adjoint_equals_per_scene(rasterizer):
    """Three scenes in one launch vs each alone: same losses and gradients
    (up to the order of atomic accumulation)."""
    items = [A.case(192, 90, 61), A.case(256, 160, 62), A.case(128, 40, 63)]
    adj = G.Adjoint(rasterizer, items)
    try:
        say("batched adjoint")
        adj.run()
        together = [adj.result(i, image=True) for i in range(3)]
        for i in range(3):
            adj.run_scene(i)
            alone = adj.result(i, image=True)
            assert abs(alone[0] - together[i][0]) <= 1e-6 * max(1.0, alone[0])
            np.testing.assert_allclose(alone[1], together[i][1], rtol=1e-4, atol=1e-6 * np.abs(alone[1]).max())
            np.testing.assert_allclose(alone[2], together[i][2], rtol=1e-4, atol=1e-6 * np.abs(alone[2]).max())
            np.testing.assert_array_equal(alone[3], together[i][3])
    finally:
        adj.close()


def test_torch_autograd_wrapper_descends(rasterizer):
    """The harness's torch.autograd.Function: backward gives the adjoint's
    gradients and a few gradient steps lower the loss."""
    import torch
    from tests.adjoint_torch import raster_l2
    sc, target = A.case(128, 40, 71)
    pairs = torch.tensor(sc.pairs.reshape(-1, 4), dtype=torch.float32, requires_grad=True)
    rgba = torch.tensor(np.asarray(sc.substances), dtype=torch.float32, requires_grad=True)
    losses = []
    opt = torch.optim.SGD([{"params": [pairs], "lr": 0.02}, {"params": [rgba], "lr": 2e-4}])
    for step in range(4):
        opt.zero_grad()
        loss = raster_l2(pairs, rgba, rasterizer, sc, target)
        loss.backward()
        losses.append(float(loss))
        say(f"step {step}: loss {losses[-1]:.4f}")
        assert torch.isfinite(pairs.grad).all() and torch.isfinite(rgba.grad).all()
        opt.step()
    assert losses[-1] < losses[0]


def _same_result(a, b, what):
    """Equal up to the order of atomic accumulation (the image bit for bit).
    A gradient is an FP32 sum of per-tile partial sums added atomically in
    whatever order the tiles finish; with cancellation among ~10^4 tiles two
    orders differ by ~1e-4 relative (a substance's colour gradient of 7.7e5
    on a 2048^2 scene: 1.09e-4), so the bound is 1e-3."""
    assert abs(a[0] - b[0]) <= 1e-6 * max(1.0, abs(a[0])), f"{what}: loss {a[0]} vs {b[0]}"
    np.testing.assert_allclose(a[1], b[1], rtol=1e-3, atol=1e-5 * max(1e-30, np.abs(a[1]).max()), err_msg=what)
    np.testing.assert_allclose(a[2], b[2], rtol=1e-3, atol=1e-5 * max(1e-30, np.abs(a[2]).max()), err_msg=what)
    np.testing.assert_array_equal(a[3], b[3], err_msg=what)


def test_device_update_equals_host_build(rasterizer):
    """Parameter updates re-serialise and re-bin on the device (with the
    curve provenance the point gradients need).  Updating a two-scene batch
    to its own parameters reproduces the host build's results; updating one
    scene to new parameters, from host memory or from a device tensor, gives
    the results of a fresh upload (host build) of those parameters."""
    import torch
    items = [A.case(192, 90, 81), A.case(256, 160, 82)]
    adj = G.Adjoint(rasterizer, items)
    fresh = None
    try:
        say("host build")
        adj.run()
        base = [adj.result(i, image=True) for i in range(2)]
        say("device rebuild, same parameters")
        adj.update(1, pairs=items[1][0].pairs)
        adj.run()
        for i in range(2):
            _same_result(base[i], adj.result(i, image=True), f"scene {i} after a same-parameter update")
        moved = A.perturbed(items[1][0], 99, point_sigma=1.5, color_sigma=0.0)
        say("device rebuild, new parameters (host pointer)")
        adj.update(1, pairs=moved.pairs)
        adj.run()
        got_host = adj.result(1, image=True)
        say("fresh upload of the new parameters (host build)")
        fresh = G.Adjoint(rasterizer, [items[0], (moved, items[1][1])])
        fresh.run()
        want = fresh.result(1, image=True)
        _same_result(want, got_host, "update from host memory")
        _same_result(base[0], fresh.result(0, image=True), "untouched scene")
        say("device rebuild, new parameters (device pointer)")
        adj.update(1, pairs=items[1][0].pairs)
        t = torch.tensor(moved.pairs.reshape(-1), dtype=torch.float32, device="cuda")
        adj.update_device(1, pairs_ptr=t.data_ptr())
        adj.run()
        dp = torch.zeros(moved.pairs.size, dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        adj.result_device(1, loss_ptr=loss.data_ptr(), d_pairs_ptr=dp.data_ptr())
        adj.check()
        got_dev = adj.result(1, image=True)
        _same_result(want, got_dev, "update from device memory")
        assert abs(float(loss.item()) - got_dev[0]) == 0.0
        np.testing.assert_array_equal(dp.cpu().numpy().reshape(got_dev[1].shape), got_dev[1])
    finally:
        adj.close()
        if fresh is not None:
            fresh.close()


def test_whole_batch_device_update(rasterizer):
    """Scene -1 updates every scene from the concatenated parameters and
    reads every scene's gradients back concatenated."""
    import torch
    items = [A.case(160, 70, 91), A.case(128, 50, 92)]
    moved = [A.perturbed(sc, 7 + i, point_sigma=1.0, color_sigma=0.0) for i, (sc, _) in enumerate(items)]
    adj = G.Adjoint(rasterizer, items)
    fresh = G.Adjoint(rasterizer, [(m, t) for m, (_, t) in zip(moved, items)])
    try:
        t = torch.tensor(np.concatenate([m.pairs.reshape(-1) for m in moved]), dtype=torch.float32, device="cuda")
        adj.update_device(-1, pairs_ptr=t.data_ptr())
        adj.run()
        fresh.run()
        dp = torch.zeros_like(t)
        adj.result_device(-1, d_pairs_ptr=dp.data_ptr())
        adj.check()
        off = 0
        for i, m in enumerate(moved):
            got, want = adj.result(i, image=True), fresh.result(i, image=True)
            _same_result(want, got, f"scene {i}")
            n = m.pairs.size
            np.testing.assert_array_equal(dp[off:off + n].cpu().numpy(), got[1].reshape(-1))
            off += n
    finally:
        adj.close()
        fresh.close()


def test_whole_batch_update_regrows_binning():
    """A whole-batch update whose one-block-per-scene binning outgrows its
    first capacities: 40 large shapes on 2048^2 at 16-pixel tiles make 16,384
    tiles and 146,895 (tile, shape) pairs against a first capacity of 32,768
    pairs, beside a small scene.  After the update every scene equals a fresh
    host-built upload."""
    import torch
    from gudni_amd import scenes as S
    big = S.batch_scene(2048, 40, 31, gradient_share=0)
    small, small_target = A.case(128, 40, 93)
    r = G.Rasterizer(G.default_spec(max_tile_size=16))
    try:
        items = [(A.perturbed(big, 5), np.zeros((big.height, big.width, 4), np.float32)), (small, small_target)]
        moved = [A.perturbed(sc, 17 + i, point_sigma=1.0, color_sigma=0.0) for i, (sc, _) in enumerate(items)]
        adj = G.Adjoint(r, items)
        fresh = G.Adjoint(r, [(m, t) for m, (_, t) in zip(moved, items)])
        try:
            t = torch.tensor(np.concatenate([m.pairs.reshape(-1) for m in moved]), dtype=torch.float32,
                             device="cuda")
            say("whole-batch update with a 2048^2 scene at 16-px tiles")
            adj.update_device(-1, pairs_ptr=t.data_ptr())
            adj.run()
            fresh.run()
            for i in range(2):
                _same_result(fresh.result(i, image=True), adj.result(i, image=True), f"scene {i}")
        finally:
            adj.close()
            fresh.close()
    finally:
        r.close()


def compound_case(size, n, seed, n_subs):
    """A blob scene whose shapes share n_subs substances and a quarter of
    which are subtract compounds (cSubtract, Figure/ShapeTree.hs:84-166):
    (perturbed scene, target)."""
    import copy
    from gudni_amd import scenes as S
    base = S.batch_scene(size, n, seed, gradient_share=0)
    base = copy.deepcopy(base)
    tags = []
    for i in range(base.n_shapes):
        tags.append(S.solid_tag(i % n_subs, subtract=(i % 4 == 1)))
    base.tags = np.array(tags, np.uint64)
    base.substances = np.ascontiguousarray(np.asarray(base.substances, np.float32)[:n_subs])
    ref = H.oracle_render(base, want_f64=False)
    target = ref["f32"].reshape(base.height, base.width, 4).astype(np.float32)
    return A.perturbed(base, seed + 1000), target


@pytest.mark.parametrize("tile", [256, 32])
def test_adjoint_compounds_and_shared_substances(tile):
    """determineColor's merge of shapes sharing a substance and its subtract
    compounds (Kernels.cl:1447-1513) in the adjoint: forward image and loss
    equal the oracle's, gradients match FP64 central differences."""
    sc, target = compound_case(256, 150, 43, 40)
    spec = G.default_spec(max_tile_size=tile)
    r = G.Rasterizer(spec)
    try:
        say(f"adjoint with compounds and shared substances, tile {tile}")
        loss, dp, dc, img = gpu_adjoint(r, sc, target)
    finally:
        r.close()
    ref = H.oracle_render(sc, spec, want_f64=True)
    err = np.abs(img.astype(np.float64) - ref["f64"].reshape(img.shape)).max()
    assert err <= COVERAGE_TOL, f"forward image differs from the oracle by {err}"
    ref_loss = float(((ref["f64"].reshape(img.shape) - target) ** 2).sum())
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    samples = A.fd_samples(sc, target, n=64, seed=9, seg=A.segment_map(sc, spec), max_tries=1200)
    assert len(samples) == 64
    got, fd, e, bad = A.compare(samples, dp, dc)
    say(f"max |adjoint - fd| {e.max():.3e}")
    assert not bad.any(), [(s, g) for s, g, b in zip(samples, got, bad) if b][:5]
