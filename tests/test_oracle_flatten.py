"""Pins the oracle's quadratic flattening -- searchTree + bifurcateCurve +
yIntercept (Kernels.cl:1226-1258, 1337-1357, 818-820; gudni_oracle.c
intersect_curve_n / search_tree) -- against a closed form computed in FP64
from the Bernstein form of the curve, with no bisection code:

  * A quadratic with end points L, R and control C has the second difference
    D = L - 2C + R.  Halving a quadratic at t = 1/2 divides its control's
    offset from the on-curve midpoint, ctl - mid = -D/4, by exactly 4, so at
    halving level n that offset is -D / 4^(n+1) and the reference's taxicab
    flatness is |D|_1 / 4^(n+1).  The loop halves while flatness > 0.25, so
    the halving count is the least n >= 0 with |D|_1 <= 4^n (0 when the
    control coincides with the left end point, which skips the loop).
  * Each halving keeps the half whose x range holds the border (xpos < mid.x
    -> left half).  On an x-monotone curve the final piece is therefore
    [k/2^n, (k+1)/2^n] with k = #{j in 1..2^n-1 : x(j/2^n) <= xpos}.
  * The ordinate is the chord intercept between the Bernstein points
    P(k/2^n) and P((k+1)/2^n) at xpos (the lower ordinate of a vertical
    chord).

The FP32 trajectory rounds every midpoint, so the oracle's ordinate is held
to a few FP32 ulps of the coordinates' magnitude (scaled by the chord's
slope), and the halving count must equal the closed form except where the
exact flatness lies within FP32 rounding of 0.25 (those cases are counted
and either count is accepted).  The same checks reject mutants of the loop
(flatness constant, branch direction, the midpoint, the intercept), built
from copies of the oracle's source on the CPU.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from gudni_amd import scenes as S
from oracle import oracle as O

EPS32 = np.finfo(np.float32).eps


def bern(L, Cc, R, t):
    """Quadratic Bernstein point in FP64 (t may be an array)."""
    t = np.asarray(t, np.float64)[..., None]
    return (1 - t) ** 2 * L + 2 * t * (1 - t) * Cc + t ** 2 * R


def closed_form(L, Cc, R, xpos, M):
    """Candidate (halvings, ordinate) pairs for the border at xpos on the
    x-monotone quadratic (L, Cc, R) (FP64 arrays of the FP32 column-local
    points).  More than one candidate only where the exact flatness lies
    within FP32 rounding of the threshold at the last level."""
    if L[0] == Cc[0] and L[1] == Cc[1]:
        ns = [0]
    else:
        a = abs(L[0] - 2 * Cc[0] + R[0]) + abs(L[1] - 2 * Cc[1] + R[1])
        n = 0
        while a > 4.0 ** n:
            n += 1
        ns = [n]
        # rounding band around the decision flat = 0.25 at levels n-1 and n
        band = (2 + 3 * n) * EPS32 * M
        if n > 0 and abs(a / 4.0 ** n - 0.25) <= band:      # level n-1 decision
            ns.append(n - 1)
        if abs(a / 4.0 ** (n + 1) - 0.25) <= band:          # level n decision
            ns.append(n + 1)
    out = []
    for n in ns:
        m = 1 << n
        xs = bern(L, Cc, R, np.arange(1, m) / m)[:, 0] if m > 1 else np.zeros(0)
        k = int(np.count_nonzero(xs <= xpos))
        k = min(k, m - 1)
        p0 = bern(L, Cc, R, k / m)
        p1 = bern(L, Cc, R, (k + 1) / m)
        if p0[0] == p1[0]:
            y = min(p0[1], p1[1])
            s = 0.0
        else:
            s = (p1[1] - p0[1]) / (p1[0] - p0[0])
            y = s * (xpos - p0[0]) + p0[1]
        out.append((n, y, s))
    return out


def strand_bytes(curves):
    """Serialise one strand of x-monotone curves [(L, C, R), ...] (consecutive
    curves share end points) in the reference's layout (Strand.hs:180-206):
    an 8-byte header (u16 size in 8-byte units), then the 2m+1 points
    permuted by the reorder table: right end point, left end point, its
    control, then the interior (point, control) nodes in the Eytzinger order
    searchTree walks."""
    m = len(curves)
    pts = []
    for c in curves:
        pts += [c[0], c[1]]
    pts.append(curves[-1][2])
    words = np.zeros(4 * (m + 1), np.float32)
    hdr = np.zeros(2, np.uint32)
    hdr[0] = 2 * m + 2
    words[0:2] = hdr.view(np.float32)
    for slot in range(2 * m + 1):
        words[2 + 2 * slot: 4 + 2 * slot] = pts[O.reorder_entry(2 * m + 1, slot)]
    return words.view(np.uint8)


def decode_strand(buf):
    """In-order curve list of a serialised strand, from its bytes: the
    interior nodes are visited by an in-order walk of the complete tree with
    children 2i+1, 2i+2 (independent of the reorder table)."""
    size = int(buf[:2].view(np.uint16)[0])
    f = buf[: 8 * size].view(np.float32)
    m = (size - 2) // 2
    right = f[2:4].copy()
    left, lctl = f[4:6].copy(), f[6:8].copy()
    nodes = f[8:].reshape(-1, 4)
    order = []

    def walk(i):
        if i < m - 1:
            walk(2 * i + 1)
            order.append(i)
            walk(2 * i + 2)
    walk(0)
    pts = [left] + [nodes[i, 0:2].copy() for i in order] + [right]
    ctls = [lctl] + [nodes[i, 2:4].copy() for i in order]
    return [(pts[i], ctls[i], pts[i + 1]) for i in range(m)]


def expected_border(curves32, dx, dy, is_left):
    """Closed-form candidates for the ordinate a strand gives the column with
    origin (dx, dy), or None if it misses the column: the curve holding the
    border is found by scanning the in-order end points (searchTree's
    left-biased <= on the left border, < on the right), then the end point
    if it lies inside the column, else the flattening closed form."""
    loc = [tuple(np.array([p[0] - np.float32(dx), p[1] - np.float32(dy)], np.float32) for p in c)
           for c in curves32]
    left0, right_end = loc[0][0], loc[-1][2]
    if not (left0[0] <= np.float32(1.0) and right_end[0] > np.float32(0.0)):
        return None
    xpos = max(np.float32(0.0), left0[0]) if is_left else min(np.float32(1.0), right_end[0])
    idx = 0
    for i in range(1, len(loc)):
        px = loc[i][0][0]
        if (xpos > px) if is_left else (xpos >= px):
            idx = i
    Lc, Cc, Rc = loc[idx]
    if is_left and Lc[0] >= 0:
        return [(-1, float(Lc[1]), 0.0)], 0.0
    if not is_left and Rc[0] <= 1:
        return [(-1, float(Rc[1]), 0.0)], 0.0
    L, Cd, R = (np.asarray(v, np.float64) for v in (Lc, Cc, Rc))
    M = float(max(np.abs(np.concatenate([L, Cd, R]))) + 1.0)
    return closed_form(L, Cd, R, float(xpos), M), M


def check(strand, curves32, dx, dy, is_left, L=None, stats=None):
    """True if the oracle's ordinate matches a closed-form candidate."""
    got = O.strand_border(strand, dx, dy, is_left, L=L)
    exp = expected_border(curves32, dx, dy, is_left)
    if got is None or exp is None:
        return (got is None) == (exp is None)
    cands, M = exp
    y, n, _ = got
    if stats is not None:
        stats["cases"] += 1
        stats["ambiguous"] += len(cands) > 1
        stats["halved"] += n > 0
    for i, (cn, cy, s) in enumerate(cands):
        if cn != n:
            continue
        if cn < 0:
            return y == np.float32(cy)
        tol = (2 + 2 * cn) * EPS32 * M * (1 + 2 * abs(s)) + 2 * EPS32 * abs(cy)
        if stats is not None:
            stats["worst"] = max(stats["worst"], abs(y - cy) / (EPS32 * M * (1 + abs(s))))
            stats["rounded"] += i > 0
        if abs(y - cy) <= tol:
            return True
    return False


def random_quadratics(seed, n):
    """x-monotone quadratics (control x between the end points) over spans of
    1..3000 px at offsets up to 4096, as single-curve strands, each with the
    integer column borders strictly inside it."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        x0 = rng.uniform(0, 4096)
        span = float(np.exp(rng.uniform(np.log(1.5), np.log(3000))))
        x1 = x0 + span
        cx = x0 + rng.uniform(0, 1) * span
        ys = rng.uniform(0, 4096) + rng.normal(0, 1, 3) * float(np.exp(rng.uniform(0, np.log(3000))))
        L = np.array([x0, ys[0]], np.float32)
        Cc = np.array([cx, ys[1]], np.float32)
        R = np.array([x1, ys[2]], np.float32)
        if not (L[0] <= Cc[0] <= R[0]) or L[0] == R[0]:
            continue
        borders = np.arange(np.ceil(L[0]), np.floor(R[0]) + 1)
        borders = borders[(borders > L[0]) & (borders < R[0])]
        if len(borders) == 0:
            continue
        pick = rng.choice(borders, size=min(6, len(borders)), replace=False)
        out.append(([(L, Cc, R)], [float(b) for b in pick]))
    return out


def run_random(L=None, n=1200, seed=7):
    stats = {"cases": 0, "ambiguous": 0, "rounded": 0, "halved": 0, "worst": 0.0}
    bad = 0
    for curves, borders in random_quadratics(seed, n):
        sb = strand_bytes(curves)
        for b in borders:
            # left border at x = b (column origin b), right border at x = b
            # (column origin b - 1): both intersect the curve
            bad += not check(sb, curves, b, 0.0, True, L, stats)
            bad += not check(sb, curves, b - 1.0, 0.0, False, L, stats)
    return bad, stats


def test_closed_form_counts_halvings():
    """The closed form's own arithmetic: |D|_1 = 4^n exactly needs n halvings,
    a hair above needs n + 1; a straight segment (control at the midpoint)
    needs none."""
    L = np.array([0.0, 0.0]); R = np.array([2.0, 0.0])
    for n in range(0, 6):
        Cc = np.array([1.0, -(4.0 ** n) / 2])          # |D|_1 = 4^n
        assert closed_form(L, Cc, R, 0.5, 1e-30)[0][0] == n
        Cc = np.array([1.0, -(4.0 ** n) / 2 * (1 + 1e-9)])
        assert closed_form(L, Cc, R, 0.5, 1e-30)[0][0] == n + 1
    assert closed_form(L, np.array([1.0, 0.0]), R, 0.5, 1.0)[0][:2] == (0, 0.0)


def test_random_quadratics_match_closed_form():
    bad, st = run_random()
    print(f"{st['cases']} border ordinates; {st['ambiguous']} with the exact flatness within "
          f"FP32 rounding of 0.25, of which {st['rounded']} took the other halving count; "
          f"worst error {st['worst']:.2f} ulp(M)(1+|s|)")
    assert st["cases"] >= 2000 and st["halved"] > 0.9 * st["cases"]
    assert bad == 0
    # the halving count is the closed form's except in the rounding band
    assert st["rounded"] <= st["ambiguous"] < 0.05 * st["cases"]
    assert st["rounded"] < 0.01 * st["cases"]


def _scene_strands(scene):
    g = O.Geometry(scene)
    shapes = g.shapes.view(np.uint32).reshape(-1, 4)
    out = []
    for rec in shapes:
        off = 16 * int(rec[2])
        for _ in range(int(rec[3])):
            size = int(g.geo[off:off + 2].view(np.uint16)[0])
            buf = g.geo[off: off + 8 * size].copy()
            out.append((buf, decode_strand(buf)))
            off += 8 * size
    return out


@pytest.mark.parametrize("name", ["simpleKnob", "blobs_scaled"])
def test_scene_strands_match_closed_form(name):
    """Every strand of the scene against every column it meets (and a random
    row origin): the ordinate on both borders of each column."""
    if name == "simpleKnob":
        pairs = np.array([[0, 0, 1, 1], [0, 2, 0, 1]], np.float32) * 100 + 100
        scene = S.build_scene(320, 320, [(S.solid_tag(0), [pairs])], [[0.25, 0.25, 0.25, 1.0]])
    else:
        scene = S.blobs_scaled(512, 400, seed=11)
    strands = _scene_strands(scene)
    rng = np.random.default_rng(3)
    stats = {"cases": 0, "ambiguous": 0, "rounded": 0, "halved": 0, "worst": 0.0}
    bad = []
    for buf, curves in strands:
        x0 = int(np.floor(curves[0][0][0])) - 1
        x1 = int(np.ceil(curves[-1][2][0])) + 1
        dy = float(rng.integers(0, 512))
        for col in range(x0, x1 + 1):
            for is_left in (True, False):
                if not check(buf, curves, float(col), dy, is_left, None, stats):
                    bad.append((col, is_left))
    print(f"{name}: {len(strands)} strands, {stats['cases']} border ordinates, "
          f"{stats['ambiguous']} in the rounding band, {stats['rounded']} took the other count, "
          f"worst {stats['worst']:.2f}")
    assert stats["cases"] > (50 if name == "simpleKnob" else 5000)
    assert stats["halved"] > 0.3 * stats["cases"], "too few ordinates from the bisection"
    assert not bad, bad[:5]


def test_strand_bytes_roundtrip():
    """The test's own strand encoder and decoder agree with the oracle's
    serialiser: a serialised scene decodes to curves that re-encode to the
    same bytes."""
    for buf, curves in _scene_strands(S.blobs_scaled(256, 60, seed=5))[:200]:
        again = strand_bytes(curves)
        assert np.array_equal(again, buf), "strand layout differs"


MUTANTS = {
    "flatness constant 0.25 -> 0.5": ("#define O_FLATNESS         0.25f", "#define O_FLATNESS         0.5f"),
    "branch direction": ("if (t.xpos < on.x) { t.ctl = lm; t.right = on; }",
                         "if (t.xpos > on.x) { t.ctl = lm; t.right = on; }"),
    "midpoint drops the control": ("v2 on = mid_point(lm, rm);", "v2 on = mid_point(t.left, t.right);"),
    "intercept from the wrong end": ("* (t.xpos - t.left.x)) + t.left.y;", "* (t.xpos - t.right.x)) + t.left.y;"),
    "euclidean flatness": ("float flat = fabsf(t.ctl.x - on.x) + fabsf(t.ctl.y - on.y);",
                           "float flat = sqrtf((t.ctl.x - on.x) * (t.ctl.x - on.x) + (t.ctl.y - on.y) * (t.ctl.y - on.y));"),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_rejected(name, tmp_path):
    """A plausible mistake planted in a copy of the oracle's loop fails the
    closed-form check."""
    src = open(os.path.join(os.path.dirname(O.LIB_PATH), "gudni_oracle.c")).read()
    old, new = MUTANTS[name]
    assert src.count(old) == 1, f"mutation site not found: {old}"
    path = tmp_path / "mutant.c"
    path.write_text(src.replace(old, new))
    inc = os.path.dirname(O.LIB_PATH)
    so = tmp_path / f"libmutant_{abs(hash(name))}.so"
    subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared", "-I", inc,
                    "-o", str(so), str(path), "-lm", "-lpthread"], check=True)
    L = C.CDLL(str(so))
    L.go_strand_border.restype = C.c_int
    L.go_strand_border.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int,
                                   C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    bad, st = run_random(L=L, n=300)
    print(f"mutant '{name}': {bad} of {st['cases']} ordinates rejected")
    assert bad > 0.05 * st["cases"], f"mutant '{name}' passed {st['cases'] - bad}/{st['cases']}"
