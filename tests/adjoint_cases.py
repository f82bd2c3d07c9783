"""Scenes, targets and finite differences for the adjoint's parity tests.

The target is the oracle's render of a seeded scene; the parameters are a
perturbed copy (every outline point + N(0, 4 px), colours + N(0, 0.1)).
Finite differences are central differences of the FP64 forward
(oracle/gudni_diff.c) over the window the parameter can affect, rejected when
any discrete decision differs between the two evaluations and the base."""
from __future__ import annotations

import copy

import numpy as np

from gudni_amd import scenes as S
from oracle import oracle as O
from tests import helpers as H


def perturbed(scene, seed, point_sigma=4.0, color_sigma=0.1):
    rng = np.random.default_rng(seed)
    sc = copy.deepcopy(scene)
    sc.pairs = (sc.pairs.astype(np.float64) + rng.normal(0.0, point_sigma, sc.pairs.shape)).astype(np.float32)
    subs = np.asarray(sc.substances, np.float64)
    subs = np.clip(subs + rng.normal(0.0, color_sigma, subs.shape), 0.0, 1.0)
    sc.substances = subs.astype(np.float32)
    return sc


def case(size, n_shapes, seed):
    """(perturbed scene, target (H, W, 4) float32)."""
    base = S.batch_scene(size, n_shapes, seed, gradient_share=0)
    ref = H.oracle_render(base, want_f64=False)
    target = ref["f32"].reshape(base.height, base.width, 4).astype(np.float32)
    return perturbed(base, seed + 1000), target


def pair_shape(scene):
    """Shape index of every outline pair."""
    owner = np.zeros(scene.pairs.size // 4, np.int64)
    so, oo = scene.shape_outline_offsets, scene.outline_pair_offsets
    for s in range(scene.n_shapes):
        owner[oo[so[s]]:oo[so[s + 1]]] = s
    return owner


def shape_window(scene, s, pad=3):
    so, oo = scene.shape_outline_offsets, scene.outline_pair_offsets
    pts = scene.pairs.reshape(-1, 4)[oo[so[s]]:oo[so[s + 1]]].reshape(-1, 2)
    x0 = int(np.floor(pts[:, 0].min())) - pad
    y0 = int(np.floor(pts[:, 1].min())) - pad
    x1 = int(np.ceil(pts[:, 0].max())) + pad + 1
    y1 = int(np.ceil(pts[:, 1].max())) + pad + 1
    return (max(0, x0), max(0, y0), min(scene.width, x1), min(scene.height, y1))


def segment_rows(scene, spec) -> int:
    """Rows per column segment of the tile tree `spec` builds for the scene:
    a tile 2^hd wide and 2^vd tall is rendered by threads_per_tile threads,
    2^(log2 threads - hd) of them stacked in each column (initTileState,
    Kernels.cl:1692-1722).  Each segment trims the records crossing its top
    and starts its own sort, and the reference's sweep does not reorder
    active records that cross inside a span, so the forward (and its
    gradient) depends on where columns are cut.  The cases here use one
    segment height everywhere; anything else is an error."""
    import gudni_amd as G
    _, _, jobs = G.prepare(scene, spec)
    t = np.frombuffer(jobs.tiles.tobytes(), np.int32).reshape(-1, 8)
    hd = (t[:, 4] & 0xFFFF).astype(np.int64)
    vd = (t[:, 4] >> 16).astype(np.int64)
    log_threads = int(spec.threads_per_tile).bit_length() - 1
    rows = {max(0, int(v) - max(0, log_threads - int(h))) for h, v in zip(hd, vd)}
    if len(rows) != 1:
        raise ValueError(f"column segments of several heights: {sorted(1 << r for r in rows)}")
    return 1 << rows.pop()


def segmented_loss(scene, pairs, rgba, target, window, seg=None):
    """The FP64 forward's loss over a window with the columns cut into
    segments of `seg` rows aligned to multiples of seg (None: one segment
    per window), and the decision signature of every segment."""
    if seg is None:
        return O.diff_loss(scene, pairs, rgba, target, window)
    x0, y0, x1, y1 = window
    y0 = (y0 // seg) * seg
    y1 = min(scene.height, -(-y1 // seg) * seg)
    return O.diff_loss(scene, pairs, rgba, target, (x0, y0, x1, y1), seg=seg)


def fd_samples(scene, target, n=64, seed=7, h_point=1e-2, h_color=1e-3, max_tries=600, seg=None):
    """n accepted samples: (kind, index, component, fd) with kind 'point'
    (index = pair, component 0..3) or 'rgba' (index = substance).  A sample
    is rejected when the decision signature at +h or -h differs from the
    base's, or when the central differences at h and h/2 disagree (a kink
    inside the step that the signature does not name).  seg: the column
    segment height of the forward being checked (segment_rows)."""
    rng = np.random.default_rng(seed)
    owner = pair_shape(scene)
    pairs = scene.pairs.reshape(-1, 4).astype(np.float64)
    rgba = np.asarray(scene.substances, np.float64)
    sid = (np.asarray(scene.tags, np.uint64) & np.uint64(0x07FFFFFFFFFFFFFF)).astype(np.int64)
    out, tries = [], 0
    while len(out) < n and tries < max_tries:
        tries += 1
        if rng.uniform() < 0.75:
            kind, p, comp = "point", int(rng.integers(len(pairs))), int(rng.integers(4))
            s, h = int(owner[p]), h_point
        else:
            s = int(rng.integers(scene.n_shapes))
            kind, p, comp, h = "rgba", int(sid[s]), int(rng.integers(4)), h_color
        win = shape_window(scene, s)
        if win[2] <= win[0] or win[3] <= win[1]:
            continue

        def loss(delta):
            pp, cc = pairs, rgba
            if kind == "point":
                pp = pairs.copy()
                pp[p, comp] += delta
            else:
                cc = rgba.copy()
                cc[p, comp] += delta
            return segmented_loss(scene, pp, cc, target, win, seg)

        l0, _, s0 = loss(0.0)
        lp, _, sp = loss(h)
        lm, _, sm = loss(-h)
        if sp != s0 or sm != s0:
            continue
        fd = (lp - lm) / (2.0 * h)
        lp2, _, _ = loss(h / 2)
        lm2, _, _ = loss(-h / 2)
        fd2 = (lp2 - lm2) / h
        if abs(fd - fd2) > 1e-3 * max(abs(fd), abs(fd2)) + 1e-12:
            continue
        out.append((kind, p, comp, fd))
    return out


def compare(samples, d_pairs, d_rgba, rel=1e-2, floor=1e-4):
    """Per sample |adjoint - fd| against rel*|fd| + floor*rms(fd of its kind)."""
    got = np.array([d_pairs[i, k] if kind == "point" else d_rgba[i, k] for kind, i, k, _ in samples], np.float64)
    fd = np.array([v for *_, v in samples])
    is_pt = np.array([kind == "point" for kind, *_ in samples])
    scale = np.zeros_like(fd)
    for sel in (is_pt, ~is_pt):
        if sel.any():
            scale[sel] = np.sqrt(np.mean(fd[sel] ** 2))
    err = np.abs(got - fd)
    bad = err > rel * np.abs(fd) + floor * scale
    return got, fd, err, bad
