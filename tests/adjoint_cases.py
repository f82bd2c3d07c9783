"""Scenes, targets and finite differences for the adjoint's parity tests.

The target is the oracle's render of a seeded scene; the parameters are a
perturbed copy (every outline point + N(0, 4 px), colours + N(0, 0.1)).
Finite differences are central differences of the FP64 forward
(oracle/gudni_diff.c) over the window the parameter can affect, rejected when
any discrete decision differs between the two evaluations and the base."""
from __future__ import annotations

import copy
import os
from concurrent.futures import ThreadPoolExecutor

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


def segment_map(scene, spec) -> np.ndarray:
    """(H, W) int32: the height of the column segment each pixel belongs to
    in the tile tree `spec` builds for the scene (2^dd of its tile,
    initTileState, Kernels.cl:1692-1722), for trees of mixed tile sizes."""
    import gudni_amd as G
    _, _, jobs = G.prepare(scene, spec)
    t = np.frombuffer(jobs.tiles.tobytes(), np.int32).reshape(-1, 8)
    hd = (t[:, 4] & 0xFFFF).astype(np.int64)
    vd = (t[:, 4] >> 16).astype(np.int64)
    log_threads = int(spec.threads_per_tile).bit_length() - 1
    out = np.zeros((scene.height, scene.width), np.int32)
    for (l, tp, r, b), h, v in zip(t[:, :4], hd, vd):
        dd = max(0, int(v) - max(0, log_threads - int(h)))
        out[max(0, tp):min(scene.height, b), max(0, l):min(scene.width, r)] = 1 << dd
    assert (out > 0).all(), "the tiles do not cover the canvas"
    return out


def has_pictures(scene) -> bool:
    t = np.asarray(scene.tags, np.uint64) & np.uint64(0xC000000000000000)
    return bool((t == np.uint64(S.TAG_PICTURE)).any())


def segmented_loss(scene, pairs, rgba, target, window, seg=None, filter=0):
    """The FP64 forward's loss over a window with the columns cut into
    segments: seg None = one segment per window; an int = segments of seg
    rows aligned to multiples of seg; an (H, W) array = the segment height
    map of a tile tree (segment_map).  Pictures are sampled with `filter`.
    Returns (loss, image, decision signature)."""
    if isinstance(seg, np.ndarray):
        return O.diff_loss_ex(scene, pairs, rgba, target, window, seg_map=seg, filter=filter)
    pict = has_pictures(scene)
    if seg is None:
        if pict:
            return O.diff_loss_ex(scene, pairs, rgba, target, window, filter=filter)
        return O.diff_loss(scene, pairs, rgba, target, window)
    x0, y0, x1, y1 = window
    y0 = (y0 // seg) * seg
    y1 = min(scene.height, -(-y1 // seg) * seg)
    if pict:
        return O.diff_loss_ex(scene, pairs, rgba, target, (x0, y0, x1, y1), filter=filter, seg=seg)
    return O.diff_loss(scene, pairs, rgba, target, (x0, y0, x1, y1), seg=seg)


def fd_samples(scene, target, n=64, seed=7, h_point=1e-2, h_color=1e-3, max_tries=600, seg=None,
               filter=0, point_shapes=None, p_point=0.75, threads=None):
    """n accepted samples: (kind, index, component, fd) with kind 'point'
    (index = pair, component 0..3) or 'rgba' (index = substance).  A sample
    is rejected when the decision signature at +h or -h differs from the
    base's, or when the central differences at h and h/2 disagree (a kink
    inside the step that the signature does not name).  seg: the column
    segments of the forward being checked (segment_rows or segment_map);
    filter: picture sampling; point_shapes: if given, point samples are
    drawn from these shapes' outlines only.  Candidates are drawn in a
    seeded order and evaluated in parallel threads (the FP64 forward runs
    outside the interpreter lock); the first n accepted, in draw order, are
    returned."""
    rng = np.random.default_rng(seed)
    owner = pair_shape(scene)
    pairs = scene.pairs.reshape(-1, 4).astype(np.float64)
    rgba = np.asarray(scene.substances, np.float64)
    sid = (np.asarray(scene.tags, np.uint64) & np.uint64(0x07FFFFFFFFFFFFFF)).astype(np.int64)
    solid = (np.asarray(scene.tags, np.uint64) & np.uint64(0xC000000000000000)) == np.uint64(S.TAG_SOLID)
    if point_shapes is not None:
        allowed = np.flatnonzero(np.isin(owner, np.asarray(list(point_shapes))))
    else:
        allowed = np.arange(len(pairs))
    cands = []
    for _ in range(max_tries):
        if rng.uniform() < p_point and len(allowed):
            kind, p, comp = "point", int(allowed[rng.integers(len(allowed))]), int(rng.integers(4))
            s, h = int(owner[p]), h_point
        else:
            s = int(rng.integers(scene.n_shapes))
            kind, p, comp, h = "rgba", int(sid[s]), int(rng.integers(4)), h_color
            if not solid[s]:
                continue   # a picture's colour is its texels, not a parameter
        if kind == "rgba":
            # a substance colours every shape that uses it
            ws = [shape_window(scene, t) for t in np.flatnonzero(sid == p)]
            win = (min(w[0] for w in ws), min(w[1] for w in ws), max(w[2] for w in ws), max(w[3] for w in ws))
        else:
            win = shape_window(scene, s)
        if win[2] <= win[0] or win[3] <= win[1]:
            continue
        cands.append((kind, p, comp, h, win))

    def evaluate(c):
        kind, p, comp, h, win = c

        def loss(delta):
            pp, cc = pairs, rgba
            if kind == "point":
                pp = pairs.copy()
                pp[p, comp] += delta
            else:
                cc = rgba.copy()
                cc[p, comp] += delta
            return segmented_loss(scene, pp, cc, target, win, seg, filter)

        l0, _, s0 = loss(0.0)
        lp, _, sp = loss(h)
        lm, _, sm = loss(-h)
        if sp != s0 or sm != s0:
            return None
        fd = (lp - lm) / (2.0 * h)
        lp2, _, _ = loss(h / 2)
        lm2, _, _ = loss(-h / 2)
        fd2 = (lp2 - lm2) / h
        if abs(fd - fd2) > 1e-3 * max(abs(fd), abs(fd2)) + 1e-12:
            return None
        return (kind, p, comp, fd)

    threads = threads or min(16, os.cpu_count() or 1)
    out = []
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for i in range(0, len(cands), 4 * threads):
            for r in ex.map(evaluate, cands[i:i + 4 * threads]):
                if r is not None and len(out) < n:
                    out.append(r)
            if len(out) >= n:
                break
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
