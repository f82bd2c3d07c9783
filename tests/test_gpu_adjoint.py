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


def test_adjoint_forward_matches_oracle(rasterizer, adjoint_small):
    sc, target = adjoint_small
    say("adjoint forward vs oracle")
    loss, dp, dc, img = gpu_adjoint(rasterizer, sc, target)
    ref = H.oracle_render(sc, want_f64=True)
    err = np.abs(img.astype(np.float64) - ref["f64"].reshape(img.shape)).max()
    assert err <= COVERAGE_TOL, f"forward image differs from the oracle by {err}"
    l64, _, _ = O.diff_loss(sc, sc.pairs.reshape(-1, 4), np.asarray(sc.substances, np.float64), target,
                            (0, 0, sc.width, sc.height))
    ref_loss = float(((ref["f64"].reshape(img.shape) - target) ** 2).sum())
    say(f"loss gpu {loss:.6f} oracle {ref_loss:.6f} fp64 forward {l64:.6f}")
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    assert np.all(np.isfinite(dp)) and np.all(np.isfinite(dc))


def test_adjoint_gradients_match_finite_differences(rasterizer, adjoint_small):
    sc, target = adjoint_small
    say("adjoint gradients")
    loss, dp, dc, _ = gpu_adjoint(rasterizer, sc, target)
    say("finite differences (64 samples, FP64)")
    samples = A.fd_samples(sc, target, n=64, seed=5)
    assert len(samples) == 64
    got, fd, err, bad = A.compare(samples, dp, dc)
    for (kind, i, k, v), g, e, b in zip(samples, got, err, bad):
        say(f"{kind:5s} {i:5d}.{k} fd {v:+.6e} adjoint {g:+.6e} err {e:.2e}{' BAD' if b else ''}")
    assert bad.sum() == 0, f"{bad.sum()} of 64 gradients off"
    # and the same algorithm in FP64 over every parameter
    _, dp64, dc64 = O.diff_grad(sc, sc.pairs, sc.substances, target)
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
