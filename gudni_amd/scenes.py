"""Seeded synthetic scenes for the raster path.

Every scene is the flat form of what ``buildOverScene`` walks
(Raster/Serialize.hs:148-177, 266-273): a sequence of shapes, each with its
ShapeTag (Raster/ShapeInfo.hs:89-99) and closed outlines of curve pairs
(Figure/Outline.hs:55-107), plus the substance table, picture data and
picture usages the kernels read.

Randomness is SplitMix64 (Steele, Lea & Flood 2014) so scenes are identical on
every machine; floats are ``(u >> 40) * 2**-24``.  All geometry is FP32 pixel
units.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MASK64 = (1 << 64) - 1

TAG_SOLID = 0x8000000000000000
TAG_PICTURE = 0x4000000000000000
# both type bits: a gradient substance (an extension; the reference has none)
TAG_GRADIENT = 0xC000000000000000
TAG_ADD = 0x2000000000000000
TAG_SUBTRACT = 0x3000000000000000
TAG_ID_MASK = 0x07FFFFFFFFFFFFFF
# fill rule bit (an extension; the reference fills by crossing parity only):
# the shape is filled where its winding number is non-zero
TAG_NONZERO = 0x0800000000000000

# SplitMix64 seeds of the benchmark scenes ("gud1".."gud5" as integers)
SEEDS = {"uniform_512": 0x67756431, "blobs_4096": 0x67756432, "batch_256": 0x67756433,
         "adjoint_1024": 0x67756434, "adjoint_batch_512": 0x67756435}


class SplitMix64:
    """Vectorised SplitMix64: ``next(n)`` returns n consecutive outputs."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & MASK64)

    def next(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            inc = np.uint64(0x9E3779B97F4A7C15)
            z = self.state + inc * np.arange(1, n + 1, dtype=np.uint64)
            self.state = np.uint64(self.state + inc * np.uint64(n))
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int) -> np.ndarray:
        """Floats in [0, 1) with 24 random bits, as float64."""
        return (self.next(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)

    def integers(self, lo: int, hi: int, n: int) -> np.ndarray:
        """Integers uniform over [lo, hi] (inclusive)."""
        span = hi - lo + 1
        return lo + (self.next(n) % np.uint64(span)).astype(np.int64)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller standard normals."""
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


@dataclasses.dataclass
class Scene:
    width: int
    height: int
    tags: np.ndarray                    # uint64 [n_shapes]
    shape_outline_offsets: np.ndarray   # int32 [n_shapes + 1]
    outline_pair_offsets: np.ndarray    # int32 [n_outlines + 1]
    pairs: np.ndarray                   # float32 [n_pairs, 4]
    substances: np.ndarray              # float32 [n_substances, 4] (picture: u32 id in .x bits)
    pict_data: np.ndarray               # uint8 RGBA pile
    pict_uses: np.ndarray               # uint8 [n_uses, 24] PictureUse records
    background: tuple = (1.0, 1.0, 1.0, 1.0)
    name: str = ""
    # gradient records, 28 float32 words (112 B) each: kind (i32: 0 linear,
    # 1 radial), stop count (i32), x0 y0 x1 y1, 2 unused, 4 offsets, 4 RGBA
    gradients: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 28), np.float32))

    @property
    def n_shapes(self) -> int:
        return int(self.tags.shape[0])

    @property
    def n_curves(self) -> int:
        return int(self.pairs.shape[0])


def pict_use_records(uses) -> np.ndarray:
    """PictureUse records, 24 bytes each (Figure/Picture.hs:183-199):
    float2 translate, int2 size, i32 memOffset, f32 scale."""
    dt = np.dtype([("tx", "<f4"), ("ty", "<f4"), ("w", "<i4"), ("h", "<i4"),
                   ("off", "<i4"), ("scale", "<f4")])
    arr = np.zeros(len(uses), dtype=dt)
    for i, (tx, ty, w, h, off, scale) in enumerate(uses):
        arr[i] = (tx, ty, w, h, off, scale)
    return arr.view(np.uint8).reshape(len(uses), 24) if len(uses) else np.zeros((0, 24), np.uint8)


def picture_substance(pict_id: int) -> np.ndarray:
    """Substance record of a picture (ShapeInfo.hs:144-148): u32 id, u32 0,
    8 undefined bytes (zero here)."""
    rec = np.zeros(4, dtype=np.uint32)
    rec[0] = pict_id
    return rec.view(np.float32)


def build_scene(width, height, shapes, substances, pict_data=None, pict_uses=(),
                background=(1.0, 1.0, 1.0, 1.0), name="", gradients=None):
    """shapes: list of (tag, [outline, ...]) with outline an (k, 4) array of
    curve pairs (on.x, on.y, off.x, off.y)."""
    tags = np.array([t for t, _ in shapes], dtype=np.uint64)
    so = [0]
    oo = [0]
    chunks = []
    for _, outlines in shapes:
        for o in outlines:
            o = np.asarray(o, dtype=np.float32).reshape(-1, 4)
            chunks.append(o)
            oo.append(oo[-1] + o.shape[0])
        so.append(len(oo) - 1)
    pairs = np.concatenate(chunks) if chunks else np.zeros((0, 4), np.float32)
    subs = np.asarray(substances, dtype=np.float32).reshape(-1, 4)
    pd = np.zeros(4, np.uint8) if pict_data is None else np.asarray(pict_data, np.uint8).ravel()
    grads = np.zeros((0, 28), np.float32) if gradients is None else \
        np.ascontiguousarray(np.asarray(gradients, np.float32).reshape(-1, 28))
    return Scene(width, height, tags, np.array(so, np.int32), np.array(oo, np.int32),
                 np.ascontiguousarray(pairs), np.ascontiguousarray(subs), pd,
                 pict_use_records(list(pict_uses)), tuple(background), name, grads)


def straight_pairs(points) -> np.ndarray:
    """Closed polygon -> curve pairs; each straight segment's control point is
    its midpoint (segmentsToCurvePairs, Outline.hs:91-103, with linear's
    lerp 0.5 = 0.5*a + 0.5*b)."""
    p = np.asarray(points, dtype=np.float32).reshape(-1, 2)
    q = np.roll(p, -1, axis=0)
    mid = (np.float32(0.5) * p + np.float32(0.5) * q).astype(np.float32)
    return np.concatenate([p, mid], axis=1).astype(np.float32)


def rectangle_pairs(x0, y0, x1, y1) -> np.ndarray:
    """rectangleCurve (Layout/Draw.hs:71-78) translated to (x0, y0)."""
    return straight_pairs([(x0, y0), (x1, y0), (x1, y1), (x0, y1)])


def solid_tag(sid: int, subtract: bool = False) -> int:
    return TAG_SOLID | (TAG_SUBTRACT if subtract else TAG_ADD) | (sid & TAG_ID_MASK)


def picture_tag(sid: int) -> int:
    return TAG_PICTURE | TAG_ADD | (sid & TAG_ID_MASK)


def gradient_tag(sid: int) -> int:
    return TAG_GRADIENT | TAG_ADD | (sid & TAG_ID_MASK)


def gradient_record(kind: int, p0, p1, offsets, colors) -> np.ndarray:
    """One 112-byte gradient record: kind 0 runs linearly from p0 to p1, kind
    1 radially around p0 with radius p1[0]; 1..4 stops (offset, RGBA)."""
    r = np.zeros(28, np.float32)
    iv = r.view(np.int32)
    n = len(offsets)
    if kind not in (0, 1) or not 1 <= n <= 4:
        raise ValueError(f"gradient kind {kind} with {n} stops: kind 0 or 1, 1..4 stops")
    iv[0], iv[1] = int(kind), n
    r[2:4] = p0
    r[4:6] = p1
    r[8:8 + n] = offsets
    r[12:12 + 4 * n] = np.asarray(colors, np.float32).ravel()
    return r


def gradient_substance(grad_id: int) -> np.ndarray:
    s = np.zeros(4, np.float32)
    s.view(np.uint32)[0] = grad_id
    return s


def random_gradients(rng: "SplitMix64", centre, half) -> np.ndarray:
    """Linear or radial gradients with 2-4 stops over boxes (centre, half size)."""
    n = len(centre)
    out = np.zeros((n, 28), np.float32)
    kind = (rng.next(n) & np.uint64(1)).astype(np.int64)
    stops = rng.integers(2, 4, n)
    ang = rng.uniform(n) * (2.0 * math.pi)
    u = rng.uniform(4 * n).reshape(n, 4)
    cols = (rng.next(16 * n) & np.uint64(0xFF)).astype(np.float32).reshape(n, 4, 4) / 255.0
    for i in range(n):
        cx, cy = float(centre[i][0]), float(centre[i][1])
        h = max(float(half[i]), 0.5)
        m = int(stops[i])
        offs = np.sort(np.concatenate([[0.0, 1.0], u[i, :m - 2]]))[:m]
        offs[-1] = 1.0
        if kind[i] == 0:
            dx, dy = math.cos(ang[i]) * h, math.sin(ang[i]) * h
            out[i] = gradient_record(0, (cx - dx, cy - dy), (cx + dx, cy + dy), offs, cols[i, :m])
        else:
            out[i] = gradient_record(1, (cx, cy), (h * (1.0 + u[i, 3]), 0.0), offs, cols[i, :m])
    return out


def _assemble(width, height, tags, k, pairs, substances, pict_data, pict_uses, name,
              background=(1.0, 1.0, 1.0, 1.0), gradients=None):
    """One outline per shape with k[i] pairs each."""
    n = len(k)
    oo = np.zeros(n + 1, np.int32)
    oo[1:] = np.cumsum(k)
    so = np.arange(n + 1, dtype=np.int32)
    grads = np.zeros((0, 28), np.float32) if gradients is None else gradients
    return Scene(width, height, np.asarray(tags, np.uint64), so, oo,
                 np.ascontiguousarray(pairs, dtype=np.float32),
                 np.ascontiguousarray(substances, dtype=np.float32),
                 pict_data, pict_uses, background, name, grads)


def random_colors(rng: SplitMix64, n: int) -> np.ndarray:
    """Random RGBA8 colours as floats byte/255 (Figure/Color.hs stores floats)."""
    b = (rng.next(n * 4) & np.uint64(0xFF)).astype(np.float32).reshape(n, 4)
    return (b / np.float32(255.0)).astype(np.float32)


def nonzero(tag: int) -> int:
    """The same shape tag with the nonzero fill rule."""
    return int(tag) | TAG_NONZERO


def uniform_512(n_shapes: int = 1000, size: int = 512, seed: int = SEEDS["uniform_512"],
                fill: str = "half") -> Scene:
    """512x512, 1000 closed shapes of 3-12 quadratics whose on- and off-curve
    points are uniform over the canvas; solid random RGBA8 colours.
    fill="half": shapes with even index fill even-odd (the reference's only
    rule, Kernels.cl:298-304), odd ones nonzero; fill="evenodd": all even-odd
    (what the reference itself can render)."""
    if fill not in ("half", "evenodd"):
        raise ValueError("fill must be 'half' or 'evenodd'")
    rng = SplitMix64(seed)
    k = rng.integers(3, 12, n_shapes)
    total = int(k.sum())
    pts = (rng.uniform(total * 4) * size).astype(np.float32).reshape(total, 4)
    colors = random_colors(rng, n_shapes)
    tags = [solid_tag(i) if (fill == "evenodd" or i % 2 == 0) else nonzero(solid_tag(i))
            for i in range(n_shapes)]
    return _assemble(size, size, tags, k, pts, colors, np.zeros(4, np.uint8),
                     np.zeros((0, 24), np.uint8), "uniform_512" if fill == "half" else "uniform_512_evenodd")


def noise_texture(rng: SplitMix64, side: int = 1024) -> np.ndarray:
    return (rng.next(side * side) & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.uint8)


def _blob_pairs(rng, cx, cy, r, k):
    """Closed blobs: k on-curve points at angles 2*pi*j/k, radius
    r*(1 + 0.25*U[-1,1]); controls at the half angles, pushed out by
    1/cos(pi/k) so the outline bulges like a circle."""
    total = int(k.sum())
    owner = np.repeat(np.arange(len(k)), k)
    start = np.concatenate([[0], np.cumsum(k)[:-1]])
    j = np.arange(total) - np.repeat(start, k)
    kk = np.repeat(k, k).astype(np.float64)
    ang = 2.0 * math.pi * j / kk
    rr = np.repeat(r, k) * (1.0 + 0.25 * (2.0 * rng.uniform(total) - 1.0))
    rc = np.repeat(r, k) * (1.0 + 0.25 * (2.0 * rng.uniform(total) - 1.0)) / np.cos(math.pi / kk)
    angc = ang + math.pi / kk
    x = np.repeat(cx, k) + rr * np.cos(ang)
    y = np.repeat(cy, k) + rr * np.sin(ang)
    xc = np.repeat(cx, k) + rc * np.cos(angc)
    yc = np.repeat(cy, k) + rc * np.sin(angc)
    return np.stack([x, y, xc, yc], axis=1).astype(np.float32), owner


def blobs_4096(n_blobs: int = 180_000, n_glyphs: int = 20_000, size: int = 4096,
               seed: int = SEEDS["blobs_4096"], texture_side: int = 1024, radius_median: float = 8.0,
               em_scale: float = 1.0, picture_share: int = 1, k_range=(4, 16),
               gradient_share: int = 2) -> Scene:
    """4096x4096, 200k shapes, ~2.1M quadratics.  Blobs: centre uniform,
    radius 8*exp(1.2*N(0,1)) (median 8 px), 4-16 quadratics.  Glyph-like
    outlines: 4-30 random quadratics inside a 12-48 px em box.  Substances (in
    tenths): picture_share a seeded 1024^2 RGBA8 noise texture (sampled as the
    spec's texture filter says: nearest is the reference's only mode),
    gradient_share linear/radial gradients of 2-4 stops over the shape's box
    (an extension; gradient_share=0 is the reference-runnable copy, which
    draws those shapes solid), the rest solid random RGBA8."""
    rng = SplitMix64(seed)
    cx = rng.uniform(n_blobs) * size
    cy = rng.uniform(n_blobs) * size
    r = radius_median * np.exp(1.2 * rng.normal(n_blobs))
    kb = rng.integers(k_range[0], k_range[1], n_blobs)
    blob_pairs, _ = _blob_pairs(rng, cx, cy, r, kb)
    kg = rng.integers(4, 30, n_glyphs)
    em = (12.0 + rng.uniform(n_glyphs) * 36.0) * em_scale
    gx = rng.uniform(n_glyphs) * size
    gy = rng.uniform(n_glyphs) * size
    tg = int(kg.sum())
    u = rng.uniform(tg * 4).reshape(tg, 4)
    gem = np.repeat(em, kg)
    glyph_pairs = np.stack([np.repeat(gx, kg) + u[:, 0] * gem, np.repeat(gy, kg) + u[:, 1] * gem,
                            np.repeat(gx, kg) + u[:, 2] * gem, np.repeat(gy, kg) + u[:, 3] * gem],
                           axis=1).astype(np.float32)
    # seeded shuffle of whole shapes so blobs and glyphs interleave
    k_all = np.concatenate([kb, kg])
    pairs_all = np.concatenate([blob_pairs, glyph_pairs])
    start_all = np.concatenate([[0], np.cumsum(k_all)[:-1]])
    n = n_blobs + n_glyphs
    perm = np.argsort(rng.next(n), kind="stable")
    k = k_all[perm]
    first = np.concatenate([[0], np.cumsum(k)[:-1]])
    within = np.arange(int(k.sum())) - np.repeat(first, k)
    pairs = pairs_all[np.repeat(start_all[perm], k) + within]
    # substances
    kind = rng.next(n) % np.uint64(10)
    picture = kind >= np.uint64(10 - picture_share)
    gradient = (kind >= np.uint64(10 - picture_share - gradient_share)) & ~picture
    colors = random_colors(rng, n)
    tex = noise_texture(rng, texture_side)
    subs = colors.copy()
    uses = []
    tags = np.empty(n, np.uint64)
    pic_idx = np.nonzero(picture)[0]
    pid = np.full(n, -1, np.int64)
    pid[pic_idx] = np.arange(len(pic_idx))
    ids = np.arange(n, dtype=np.uint64)
    tags[:] = np.uint64(TAG_SOLID | TAG_ADD) | ids
    tags[picture] = np.uint64(TAG_PICTURE | TAG_ADD) | ids[picture]
    sub_u32 = subs.view(np.uint32)
    sub_u32[picture, 0] = pid[picture].astype(np.uint32)
    sub_u32[picture, 1:] = 0
    grads = None
    if gradient.any():
        gi = np.nonzero(gradient)[0]
        pts = pairs.reshape(-1, 4)
        lo = np.minimum.reduceat(np.minimum(pts[:, :2], pts[:, 2:]), first, axis=0)[gi]
        hi = np.maximum.reduceat(np.maximum(pts[:, :2], pts[:, 2:]), first, axis=0)[gi]
        grads = random_gradients(rng, (lo + hi) * 0.5, np.max(hi - lo, axis=1) * 0.5)
        tags[gi] = np.uint64(TAG_GRADIENT | TAG_ADD) | ids[gi]
        sub_u32[gi, 0] = np.arange(len(gi), dtype=np.uint32)
        sub_u32[gi, 1:] = 0
    # each picture usage maps the texture's origin onto a random point of the canvas
    ox = rng.uniform(len(pic_idx)) * size
    oy = rng.uniform(len(pic_idx)) * size
    for j in range(len(pic_idx)):
        uses.append((-float(np.float32(ox[j])), -float(np.float32(oy[j])), texture_side,
                     texture_side, 0, 1.0))
    return _assemble(size, size, tags, k, pairs, subs, tex, pict_use_records(uses), "blobs",
                     gradients=grads)


def blobs_scaled(size: int, n_shapes: int, seed: int, glyph_share: float = 0.1) -> Scene:
    """The 4096^2 blob generator at another canvas size and shape count."""
    n_g = int(round(n_shapes * glyph_share))
    return blobs_4096(n_blobs=n_shapes - n_g, n_glyphs=max(n_g, 1), size=size, seed=seed,
                      texture_side=256)


def batch_scene(size: int, n_shapes: int, seed: int, k_range=(3, 12), gradient_share: int = 2,
                picture_share: int = 0, texture_side: int = 16) -> Scene:
    """One scene of the batched workload: blobs and glyph
    outlines on a size x size canvas, radii scaled so that the expected
    overdraw matches the 4096^2 scene's; solid and (gradient_share tenths)
    gradient substances."""
    radius = 8.0 * (size / 4096.0) * math.sqrt(200_000.0 / max(n_shapes, 1))
    radius = float(min(max(radius, 1.5), size / 8.0))
    n_g = max(1, n_shapes // 10)
    sc = blobs_4096(n_blobs=max(1, n_shapes - n_g), n_glyphs=n_g, size=size, seed=seed,
                    texture_side=texture_side, radius_median=radius, em_scale=max(0.25, radius / 8.0),
                    picture_share=picture_share, k_range=k_range, gradient_share=gradient_share)
    sc.name = f"batch{size}x{n_shapes}"
    return sc


def batch_256(n_scenes: int = 256, seed: int = SEEDS["batch_256"]):
    """256 independent scenes: side uniform over 256..2048, shape count
    log-uniform over [100, 50000], 3-12 quadratics per blob."""
    rng = SplitMix64(seed)
    sides = rng.integers(256, 2048, n_scenes)
    counts = np.exp(np.log(100.0) + rng.uniform(n_scenes) * (np.log(50_000.0) - np.log(100.0)))
    seeds = rng.next(n_scenes)
    return [batch_scene(int(s), int(c), int(z & np.uint64(0x7FFFFFFF)))
            for s, c, z in zip(sides, counts, seeds)]


def perturb(scene: Scene, seed: int, point_sigma: float = 4.0, color_sigma: float = 0.1) -> Scene:
    """A copy whose outline points move by N(0, point_sigma) px and whose solid
    colours move by N(0, color_sigma), clipped to [0, 1] (the learnable
    parameters of the differentiable workloads)."""
    rng = SplitMix64(seed)
    pairs = scene.pairs.astype(np.float64) + point_sigma * rng.normal(scene.pairs.size).reshape(scene.pairs.shape)
    subs = np.array(scene.substances, np.float32, copy=True)
    solid = (np.asarray(scene.tags, np.uint64) & np.uint64(0xC000000000000000)) == np.uint64(TAG_SOLID)
    sid = (np.asarray(scene.tags, np.uint64) & np.uint64(TAG_ID_MASK)).astype(np.int64)
    ids = np.unique(sid[solid])
    noise = color_sigma * rng.normal(4 * len(ids)).reshape(-1, 4)
    subs[ids] = np.clip(subs[ids].astype(np.float64) + noise, 0.0, 1.0).astype(np.float32)
    return dataclasses.replace(scene, pairs=pairs.astype(np.float32), substances=subs,
                               name=scene.name + "-perturbed")


def adjoint_1024(seed: int = SEEDS["adjoint_1024"]) -> Scene:
    """Differentiable render: a 1024x1024 scene of 2,000 shapes (the blob
    generator scaled, solid substances); the target is its render and the
    parameters a perturbed copy (``perturb``)."""
    sc = batch_scene(1024, 2000, seed & 0x7FFFFFFF, gradient_share=0)
    sc.name = "adjoint_1024"
    return sc


def adjoint_batch_512(n_scenes: int = 64, seed: int = SEEDS["adjoint_batch_512"]):
    """Batched differentiable render: 64 scenes of 512x512 with 500-5,000
    shapes (log-uniform) of 3-8 quadratics, solid plus a 1024^2 RGBA8 noise
    texture (sampled with the spec's texture filter)."""
    rng = SplitMix64(seed)
    counts = np.exp(np.log(500.0) + rng.uniform(n_scenes) * (np.log(5000.0) - np.log(500.0)))
    seeds = rng.next(n_scenes)
    out = []
    for c, z in zip(counts, seeds):
        sc = batch_scene(512, int(c), int(z & np.uint64(0x7FFFFFFF)), k_range=(3, 8), gradient_share=0,
                         picture_share=1, texture_side=1024)
        sc.name = f"adjoint-512-{int(c)}"
        out.append(sc)
    return out
