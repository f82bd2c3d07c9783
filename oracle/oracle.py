"""ctypes wrapper of the oracle (CPU restatement of Gudni's raster path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only to check or time the reference
algorithm.  The gudni_amd product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_gudni.so")


class Spec(C.Structure):
    _fields_ = [("max_thresholds", C.c_int32), ("max_shapes", C.c_int32),
                ("compute_depth", C.c_int32), ("threads_per_tile", C.c_int32),
                ("tiles_per_call", C.c_int32), ("max_tile_size", C.c_int32),
                ("max_strands_per_tile", C.c_int32), ("random_field_size", C.c_int32),
                ("stochastic_factor", C.c_float), ("coverage_mode", C.c_int32),
                ("texture_filter", C.c_int32)]


class Serialized(C.Structure):
    _fields_ = [("geo", C.POINTER(C.c_uint8)), ("geo_bytes", C.c_long),
                ("shapes", C.POINTER(C.c_uint8)), ("bbox", C.POINTER(C.c_float)),
                ("strand_count", C.POINTER(C.c_uint32)),
                ("source_index", C.POINTER(C.c_int32)), ("n_shapes", C.c_int32)]


class Jobs(C.Structure):
    _fields_ = [("n_jobs", C.c_int32), ("job_tile_offset", C.POINTER(C.c_int32)),
                ("job_shape_offset", C.POINTER(C.c_int32)), ("tiles", C.POINTER(C.c_uint8)),
                ("shape_refs", C.POINTER(C.c_uint8)), ("n_tiles", C.c_int32),
                ("n_refs", C.c_int32)]


class JobIn(C.Structure):
    _fields_ = [("geo", C.c_void_p), ("shapes", C.c_void_p), ("tiles", C.c_void_p),
                ("n_tiles", C.c_int32), ("substances", C.c_void_p), ("n_substances", C.c_int32),
                ("pict_data", C.c_void_p), ("pict_bytes", C.c_long), ("pict_uses", C.c_void_p),
                ("n_pict_uses", C.c_int32), ("random_field", C.c_void_p),
                ("background", C.c_float * 4), ("width", C.c_int32), ("height", C.c_int32),
                ("frame", C.c_int32), ("job_index", C.c_int32), ("gradients", C.c_void_p),
                ("n_gradients", C.c_int32)]


class Dump(C.Structure):
    _fields_ = [("cap", C.c_int32), ("pre_thresholds", C.c_void_p), ("pre_headers", C.c_void_p),
                ("thresholds", C.c_void_p), ("headers", C.c_void_p), ("queue_len", C.c_void_p),
                ("shape_bits", C.c_void_p), ("shape_stack", C.c_void_p),
                ("shape_indices", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        from . import build as _build
        if _build.stale():      # never check against a stale build of the oracle
            _build.build()
        L = C.CDLL(LIB_PATH)
        L.go_reorder_entry.restype = C.c_int
        L.go_reorder_entry.argtypes = [C.c_int, C.c_int]
        L.go_fix_knob.restype = C.c_int
        L.go_fix_knob.argtypes = [C.c_void_p, C.c_void_p]
        L.go_serialize_scene.restype = C.c_int
        L.go_serialize_scene.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_float, C.c_float, C.POINTER(Serialized)]
        L.go_free_serialized.argtypes = [C.POINTER(Serialized)]
        L.go_build_jobs.restype = C.c_int
        L.go_build_jobs.argtypes = [C.POINTER(Spec), C.c_int32, C.c_int32,
                                    C.POINTER(Serialized), C.POINTER(Jobs)]
        L.go_free_jobs.argtypes = [C.POINTER(Jobs)]
        L.go_raster_job.restype = C.c_int
        L.go_raster_job.argtypes = [C.POINTER(Spec), C.POINTER(JobIn), C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.POINTER(Dump)]
        L.go_render_scene.restype = C.c_int
        L.go_render_scene.argtypes = [C.POINTER(Spec), C.POINTER(Serialized), C.POINTER(Jobs),
                                      C.POINTER(JobIn), C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int]
        L.go_default_spec.argtypes = [C.POINTER(Spec)]
        L.go_strand_border.restype = C.c_int
        L.go_strand_border.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_int,
                                       C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.go_diff_loss.restype = C.c_int
        L.go_diff_loss.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                   C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.POINTER(C.c_double), C.c_void_p, C.POINTER(C.c_uint64)]
        L.go_diff_loss_ex.restype = C.c_int
        L.go_diff_loss_ex.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                      C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                      C.c_int32, C.POINTER(C.c_double), C.c_void_p,
                                      C.POINTER(C.c_uint64)]
        L.go_diff_grad.restype = C.c_int
        L.go_diff_grad.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                   C.c_int32, C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                   C.c_int32, C.c_int32]
        _lib = L
    return _lib


def default_spec(**overrides) -> Spec:
    s = Spec()
    lib().go_default_spec(C.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def strand_border(strand: np.ndarray, dx: float, dy: float, is_left: bool, L=None):
    """Border ordinate one strand gives a column (go_strand_border):
    (y, halvings, node) in column-local coordinates, or None if the strand
    misses the column.  `strand` is the strand's bytes (header first);
    `L` selects another build of the oracle library (mutant tests)."""
    L = L or lib()
    b = np.ascontiguousarray(strand, np.uint8)
    y, n, node = C.c_float(), C.c_int(), C.c_int()
    ok = L.go_strand_border(_ptr(b), float(dx), float(dy), int(bool(is_left)),
                            C.byref(y), C.byref(n), C.byref(node))
    return (y.value, n.value, node.value) if ok else None


def reorder_entry(size: int, i: int) -> int:
    return lib().go_reorder_entry(size, i)


def fix_knob(bez6) -> np.ndarray:
    b = np.ascontiguousarray(bez6, dtype=np.float32)
    o = np.zeros(12, np.float32)
    n = lib().go_fix_knob(_ptr(b), _ptr(o))
    return o[: 6 * n].reshape(n, 6)


class Geometry:
    """Serialised scene (geometry heap, shape records, bbox, strand counts)."""

    def __init__(self, scene):
        self.s = Serialized()
        self._keep = [np.ascontiguousarray(scene.tags, np.uint64),
                      np.ascontiguousarray(scene.shape_outline_offsets, np.int32),
                      np.ascontiguousarray(scene.outline_pair_offsets, np.int32),
                      np.ascontiguousarray(scene.pairs, np.float32)]
        rc = lib().go_serialize_scene(scene.n_shapes, *[_ptr(a) for a in self._keep],
                                      float(scene.width), float(scene.height), C.byref(self.s))
        if rc != 0:
            raise RuntimeError(f"go_serialize_scene failed: {rc}")
        n = self.s.n_shapes
        self.geo = np.ctypeslib.as_array(self.s.geo, (max(self.s.geo_bytes, 1),))[: self.s.geo_bytes].copy() \
            if self.s.geo_bytes else np.zeros(0, np.uint8)
        self.shapes = np.ctypeslib.as_array(self.s.shapes, (n * 16,)).copy() if n else np.zeros(0, np.uint8)
        self.bbox = np.ctypeslib.as_array(self.s.bbox, (n * 4,)).reshape(n, 4).copy() if n else np.zeros((0, 4), np.float32)
        self.strand_count = np.ctypeslib.as_array(self.s.strand_count, (n,)).copy() if n else np.zeros(0, np.uint32)
        self.source_index = np.ctypeslib.as_array(self.s.source_index, (n,)).copy() if n else np.zeros(0, np.int32)

    def __del__(self):
        try:
            lib().go_free_serialized(C.byref(self.s))
        except Exception:
            pass


class JobList:
    """Tile tree flattened into raster jobs (literal insert/split procedure)."""

    def __init__(self, spec: Spec, geometry: Geometry, width: int, height: int):
        self.j = Jobs()
        rc = lib().go_build_jobs(C.byref(spec), width, height, C.byref(geometry.s), C.byref(self.j))
        if rc != 0:
            raise RuntimeError(f"go_build_jobs failed: {rc}")
        j = self.j
        self.n_jobs = j.n_jobs
        self.job_tile_offset = np.ctypeslib.as_array(j.job_tile_offset, (j.n_jobs + 1,)).copy()
        self.job_shape_offset = np.ctypeslib.as_array(j.job_shape_offset, (j.n_jobs + 1,)).copy()
        self.tiles = np.ctypeslib.as_array(j.tiles, (j.n_tiles * 32,)).copy()
        self.shape_refs = np.ctypeslib.as_array(j.shape_refs, (max(j.n_refs, 1) * 16,))[: j.n_refs * 16].copy() \
            if j.n_refs else np.zeros(0, np.uint8)

    def job(self, i):
        t0, t1 = self.job_tile_offset[i], self.job_tile_offset[i + 1]
        s0, s1 = self.job_shape_offset[i], self.job_shape_offset[i + 1]
        return self.tiles[t0 * 32: t1 * 32], self.shape_refs[s0 * 16: s1 * 16]

    def __del__(self):
        try:
            lib().go_free_jobs(C.byref(self.j))
        except Exception:
            pass


def random_field(n: int = 4096) -> np.ndarray:
    """Random field (Util/RandomField.hs): the values are inert while
    STOCHASTIC_FACTOR is 0; a fixed field keeps runs reproducible."""
    x = np.arange(n, dtype=np.float64)
    return ((x * 0.6180339887498949) % 1.0).astype(np.float32)


def _job_in(geo, shapes, tiles, scene, rnd, job_index):
    ji = JobIn()
    ji.geo = _ptr(geo) if geo.size else _ptr(np.zeros(16, np.uint8))
    ji.shapes = _ptr(shapes) if shapes.size else _ptr(np.zeros(16, np.uint8))
    ji.tiles = _ptr(tiles)
    ji.n_tiles = tiles.size // 32
    ji.substances = _ptr(scene.substances)
    ji.n_substances = scene.substances.shape[0]
    ji.pict_data = _ptr(scene.pict_data)
    ji.pict_bytes = scene.pict_data.size
    ji.pict_uses = _ptr(scene.pict_uses) if scene.pict_uses.size else _ptr(np.zeros(24, np.uint8))
    ji.n_pict_uses = scene.pict_uses.shape[0]
    ji.random_field = _ptr(rnd)
    for i in range(4):
        ji.background[i] = float(scene.background[i])
    ji.width, ji.height, ji.frame, ji.job_index = scene.width, scene.height, 0, job_index
    grads = getattr(scene, "gradients", None)
    if grads is not None and len(grads):
        ji.gradients = _ptr(np.ascontiguousarray(grads, np.float32))
        ji.n_gradients = len(grads)
    return ji


def render(scene, spec: Spec | None = None, geometry=None, jobs=None, want_f32=False,
           want_f64=False, nthreads: int = 1, rnd=None):
    """Serialise, bin and raster a whole scene.  Returns dict with 'out'
    (H, W) uint32 BGRA words and optionally 'f32' / 'f64' (H, W, 4)."""
    spec = spec or default_spec()
    geometry = geometry or Geometry(scene)
    jobs = jobs or JobList(spec, geometry, scene.width, scene.height)
    rnd = random_field(spec.random_field_size) if rnd is None else rnd
    W, H = scene.width, scene.height
    out = np.zeros(W * H, np.uint32)
    f32 = np.zeros(W * H * 4, np.float32) if want_f32 else None
    f64 = np.zeros(W * H * 4, np.float64) if want_f64 else None
    ji = _job_in(geometry.geo, jobs.shape_refs, jobs.tiles, scene, rnd, 0)
    rc = lib().go_render_scene(C.byref(spec), C.byref(geometry.s), C.byref(jobs.j), C.byref(ji),
                               _ptr(out), _ptr(f32), _ptr(f64), nthreads)
    res = {"out": out.reshape(H, W), "rc": rc, "geometry": geometry, "jobs": jobs}
    if want_f32:
        res["f32"] = f32.reshape(H, W, 4)
    if want_f64:
        res["f64"] = f64.reshape(H, W, 4)
    return res


def raster_job_dump(scene, spec, geometry, jobs, job: int, cap: int = 64, rnd=None):
    """Run one job with the per-column dumps; returns (rc, out, dump dict)."""
    rnd = random_field(spec.random_field_size) if rnd is None else rnd
    tiles, shapes = jobs.job(job)
    n_tiles = tiles.size // 32
    slots = n_tiles * spec.threads_per_tile
    d = {"pre_thr": np.zeros((slots, cap, 4), np.float32), "pre_hdr": np.zeros((slots, cap), np.uint32),
         "thr": np.zeros((slots, cap, 4), np.float32), "hdr": np.zeros((slots, cap), np.uint32),
         "qlen": np.full(slots, -1, np.int32), "shape_bits": np.zeros(slots, np.uint32),
         "stack": np.zeros((slots, 2), np.uint64),
         "shape_indices": np.zeros((slots, spec.max_shapes), np.int32)}
    dump = Dump(cap, _ptr(d["pre_thr"]), _ptr(d["pre_hdr"]), _ptr(d["thr"]), _ptr(d["hdr"]),
                _ptr(d["qlen"]), _ptr(d["shape_bits"]), _ptr(d["stack"]), _ptr(d["shape_indices"]))
    out = np.zeros(scene.width * scene.height, np.uint32)
    ji = _job_in(geometry.geo, shapes, tiles, scene, rnd, jobs.n_jobs - 1 - job)
    rc = lib().go_raster_job(C.byref(spec), C.byref(ji), _ptr(out), None, None, C.byref(dump))
    return rc, out.reshape(scene.height, scene.width), d


def render_jobs(scene, spec, geometry, jobs, job_ids, out, nthreads=1, rnd=None):
    """Render only the listed jobs of a JobList into out (W*H uint32).
    Returns the number of pixels covered by their tiles."""
    rnd = random_field(spec.random_field_size) if rnd is None else rnd
    pixels = 0
    for j in job_ids:
        tiles, shapes = jobs.job(j)
        t = np.ascontiguousarray(tiles)
        s = np.ascontiguousarray(shapes) if shapes.size else np.zeros(16, np.uint8)
        offs_t = np.array([0, t.size // 32], np.int32)
        offs_s = np.array([0, shapes.size // 16], np.int32)
        sub = Jobs(1, offs_t.ctypes.data_as(C.POINTER(C.c_int32)),
                   offs_s.ctypes.data_as(C.POINTER(C.c_int32)),
                   t.ctypes.data_as(C.POINTER(C.c_uint8)), s.ctypes.data_as(C.POINTER(C.c_uint8)),
                   t.size // 32, shapes.size // 16)
        ji = _job_in(geometry.geo, s, t, scene, rnd, jobs.n_jobs - 1 - j)
        rc = lib().go_render_scene(C.byref(spec), C.byref(geometry.s), C.byref(sub), C.byref(ji),
                                   _ptr(out), None, None, nthreads)
        if rc != 0:
            raise RuntimeError(f"oracle failed on job {j}: {rc}")
        tr = t.view(np.int32).reshape(-1, 8)
        w = np.clip(np.minimum(tr[:, 2], scene.width) - tr[:, 0], 0, None)
        h = np.clip(np.minimum(tr[:, 3], scene.height) - tr[:, 1], 0, None)
        pixels += int((w * h).sum())
    return pixels


def diff_loss(scene, pairs, rgba, target, window, want_image=False, seg=0):
    """FP64 forward of the differentiable path (gudni_diff.c) over a window
    (x0, y0, x1, y1) of pixel columns: (loss, image or None, signature).
    seg > 0 cuts the columns into segments of seg rows at multiples of seg
    (y0 must be one, y1 one or the canvas height); 0: one segment.
    pairs: (n_pairs, 4) outline points, rgba: (n_substances, 4) solid
    colours, target: (H, W, 4) float32."""
    x0, y0, x1, y1 = [int(v) for v in window]
    tags = np.ascontiguousarray(scene.tags, np.uint64)
    soff = np.ascontiguousarray(scene.shape_outline_offsets, np.int32)
    ooff = np.ascontiguousarray(scene.outline_pair_offsets, np.int32)
    pairs = np.ascontiguousarray(pairs, np.float64)
    rgba = np.ascontiguousarray(rgba, np.float64)
    bg = np.ascontiguousarray(scene.background, np.float64)
    target = np.ascontiguousarray(target, np.float32)
    img = np.zeros((y1 - y0, x1 - x0, 4), np.float64) if want_image else None
    loss = C.c_double()
    sig = C.c_uint64()
    rc = lib().go_diff_loss(scene.n_shapes, _ptr(tags), _ptr(soff), _ptr(ooff), _ptr(pairs),
                            _ptr(rgba), _ptr(bg), scene.width, scene.height, _ptr(target),
                            x0, y0, x1, y1, int(seg), C.byref(loss), _ptr(img), C.byref(sig))
    if rc != 0:
        raise RuntimeError(f"go_diff_loss failed ({rc})")
    return loss.value, img, sig.value


def diff_loss_ex(scene, pairs, rgba, target, window, seg_map=None, filter=0, want_image=False, seg=0):
    """diff_loss with a per-pixel segment-height map (H, W) int32 (columns
    cut as a tile tree of mixed tile sizes cuts them; each column runs over
    the whole segments holding the window's rows) and picture substances
    (the scene's pictures, sampled nearest (filter 0) or bilinear (1))."""
    x0, y0, x1, y1 = [int(v) for v in window]
    tags = np.ascontiguousarray(scene.tags, np.uint64)
    soff = np.ascontiguousarray(scene.shape_outline_offsets, np.int32)
    ooff = np.ascontiguousarray(scene.outline_pair_offsets, np.int32)
    pairs = np.ascontiguousarray(pairs, np.float64)
    rgba = np.ascontiguousarray(rgba, np.float64)
    bg = np.ascontiguousarray(scene.background, np.float64)
    target = np.ascontiguousarray(target, np.float32)
    smap = None if seg_map is None else np.ascontiguousarray(seg_map, np.int32)
    if smap is not None and smap.shape != (scene.height, scene.width):
        raise ValueError("seg_map must be (height, width)")
    subs = np.ascontiguousarray(scene.substances, np.float32)
    pict = np.ascontiguousarray(scene.pict_data, np.uint8)
    uses = np.ascontiguousarray(scene.pict_uses, np.uint8)
    img = np.zeros((y1 - y0, x1 - x0, 4), np.float64) if want_image else None
    loss = C.c_double()
    sig = C.c_uint64()
    rc = lib().go_diff_loss_ex(scene.n_shapes, _ptr(tags), _ptr(soff), _ptr(ooff), _ptr(pairs),
                               _ptr(rgba), _ptr(bg), scene.width, scene.height, _ptr(target),
                               x0, y0, x1, y1, int(seg), _ptr(smap), _ptr(subs),
                               _ptr(pict) if pict.size else None, _ptr(uses) if uses.size else None,
                               uses.size // 24, int(filter), C.byref(loss), _ptr(img), C.byref(sig))
    if rc != 0:
        raise RuntimeError(f"go_diff_loss_ex failed ({rc})")
    return loss.value, img, sig.value


def diff_grad(scene, pairs, rgba, target, seg=256):
    """The adjoint's algorithm in FP64 (gudni_diff.c): (loss, d_pairs, d_rgba)."""
    tags = np.ascontiguousarray(scene.tags, np.uint64)
    soff = np.ascontiguousarray(scene.shape_outline_offsets, np.int32)
    ooff = np.ascontiguousarray(scene.outline_pair_offsets, np.int32)
    pairs = np.ascontiguousarray(pairs, np.float64).reshape(-1, 4)
    rgba = np.ascontiguousarray(rgba, np.float64).reshape(-1, 4)
    bg = np.ascontiguousarray(scene.background, np.float64)
    target = np.ascontiguousarray(target, np.float32)
    dp = np.zeros_like(pairs)
    dc = np.zeros_like(rgba)
    loss = C.c_double()
    rc = lib().go_diff_grad(scene.n_shapes, _ptr(tags), _ptr(soff), _ptr(ooff), _ptr(pairs), _ptr(rgba),
                            _ptr(bg), scene.width, scene.height, _ptr(target), int(seg), C.byref(loss),
                            _ptr(dp), _ptr(dc), len(pairs), len(rgba))
    if rc != 0:
        raise RuntimeError(f"go_diff_grad failed ({rc})")
    return loss.value, dp, dc
