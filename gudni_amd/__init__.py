"""gudni_amd -- MI355X (gfx950) raster path of Gudni behind its own C ABI.

Python mirror of the reference's host interface for this path:

  Rasterizer(spec)          setupOpenCL (OpenCL/Setup.hs:102-147)
  RasterSpec                RasterSpec (OpenCL/Rasterizer.hs:37-51)
  Geometry(scene)           buildOverScene/onShape serialisation
                            (Raster/Serialize.hs:148-177)
  RasterJobs(spec, geo)     buildTileTree + addShapeToTree + buildRasterJobs
                            (Raster/TileTree.hs:81-190, CallKernels.hs:244-255)
  Rasterizer.raster_job     generateCall (OpenCL/CallKernels.hs:88-179)
  Rasterizer.queue_jobs     queueRasterJobs (OpenCL/CallKernels.hs:218-242)

Every device call goes through libgudni_hip.so; if it is missing the
import of the native library raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Spec as RasterSpec

__all__ = ["RasterSpec", "default_spec", "Geometry", "RasterJobs", "Rasterizer", "GudniError",
           "random_field", "Batch"]

ERRORS = {-1: "bad argument", -2: "HIP runtime error", -3: "threshold queue overflow",
          -4: "section guard exceeded", -5: "no HIP device"}


class GudniError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}): {message}")
        self.code = code


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _gradients(scene) -> np.ndarray:
    """The scene's gradient records (28 float32 words each), possibly none."""
    g = getattr(scene, "gradients", None)
    if g is None or len(g) == 0:
        return np.zeros((0, 28), np.float32)
    return np.ascontiguousarray(np.asarray(g, np.float32).reshape(-1, 28))


FORMAT_BGRA8, FORMAT_RGBA8 = 0, 1


def default_spec(**overrides) -> RasterSpec:
    s = RasterSpec()
    _lib.load().gudni_default_spec(C.byref(s))
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def random_field(n: int = 4096) -> np.ndarray:
    """The random field handed to the kernels (Util/RandomField.hs).  Its
    values only matter when stochastic_factor != 0; a fixed golden-ratio
    sequence keeps frames reproducible."""
    x = np.arange(n, dtype=np.float64)
    return ((x * 0.6180339887498949) % 1.0).astype(np.float32)


def _check_host(rc):
    if rc != 0:
        msg = _lib.load().gudni_last_error(None)
        raise GudniError(rc, msg.decode() if msg else "")


class Geometry:
    """Serialised scene: strand heap, 16-byte shape records, bounding boxes and
    strand counts of the shapes that survive excludeBox."""

    def __init__(self, scene):
        L = _lib.load()
        self._arrays = [np.ascontiguousarray(scene.tags, np.uint64),
                        np.ascontiguousarray(scene.shape_outline_offsets, np.int32),
                        np.ascontiguousarray(scene.outline_pair_offsets, np.int32),
                        np.ascontiguousarray(scene.pairs, np.float32)]
        o = _lib.Outlines(scene.n_shapes, *[a.ctypes.data for a in self._arrays])
        self.handle = C.c_void_p()
        _check_host(L.gudni_geometry_build(C.byref(o), scene.width, scene.height,
                                           C.byref(self.handle)))
        v = _lib.GeometryView()
        L.gudni_geometry_view_get(self.handle, C.byref(v))
        n = v.n_shapes
        self.n_shapes = n
        self.geo = np.ctypeslib.as_array(v.geo, (v.geo_bytes,)).copy() if v.geo_bytes else np.zeros(0, np.uint8)
        self.shapes = np.ctypeslib.as_array(v.shapes, (n * 16,)).copy() if n else np.zeros(0, np.uint8)
        self.bbox = np.ctypeslib.as_array(v.bbox, (n * 4,)).reshape(n, 4).copy() if n else np.zeros((0, 4), np.float32)
        self.strand_count = np.ctypeslib.as_array(v.strand_count, (n,)).copy() if n else np.zeros(0, np.uint32)
        self.source_index = np.ctypeslib.as_array(v.source_index, (n,)).copy() if n else np.zeros(0, np.int32)

    def __del__(self):
        try:
            if self.handle:
                _lib.load().gudni_geometry_free(self.handle)
                self.handle = None
        except Exception:
            pass


class RasterJobs:
    """The tile tree flattened into raster jobs, in traversal order."""

    def __init__(self, spec: RasterSpec, geometry: Geometry, width: int, height: int):
        L = _lib.load()
        self.handle = C.c_void_p()
        _check_host(L.gudni_jobs_build(C.byref(spec), geometry.handle, width, height,
                                       C.byref(self.handle)))
        v = _lib.JobsView()
        L.gudni_jobs_view_get(self.handle, C.byref(v))
        self.n_jobs = v.n_jobs
        self.n_tiles = v.n_tiles
        self.n_refs = v.n_refs
        self.job_tile_offset = np.ctypeslib.as_array(v.job_tile_offset, (v.n_jobs + 1,)).copy()
        self.job_shape_offset = np.ctypeslib.as_array(v.job_shape_offset, (v.n_jobs + 1,)).copy()
        self.tiles = np.ctypeslib.as_array(v.tiles, (v.n_tiles * 32,)).copy()
        self.shape_refs = np.ctypeslib.as_array(v.shape_refs, (v.n_refs * 16,)).copy() \
            if v.n_refs else np.zeros(0, np.uint8)

    def job(self, i):
        t0, t1 = self.job_tile_offset[i], self.job_tile_offset[i + 1]
        s0, s1 = self.job_shape_offset[i], self.job_shape_offset[i + 1]
        return self.tiles[t0 * 32: t1 * 32], self.shape_refs[s0 * 16: s1 * 16]

    def __del__(self):
        try:
            if self.handle:
                _lib.load().gudni_jobs_free(self.handle)
                self.handle = None
        except Exception:
            pass


class Rasterizer:
    """One device context (the reference's Rasterizer)."""

    def __init__(self, spec: RasterSpec | None = None, device: int = 0):
        self.L = _lib.load()
        self.spec = spec or default_spec()
        self.ctx = C.c_void_p()
        rc = self.L.gudni_create(C.byref(self.spec), device, C.byref(self.ctx))
        if rc != 0:
            msg = self.L.gudni_last_error(None)
            raise GudniError(rc, msg.decode() if msg else "")
        self._keep = []

    def _check(self, rc):
        if rc != 0:
            msg = self.L.gudni_last_error(self.ctx)
            raise GudniError(rc, msg.decode() if msg else "")

    def close(self):
        if self.ctx:
            self.L.gudni_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int | None):
        self._check(self.L.gudni_set_stream(self.ctx, C.c_void_p(stream_handle or 0)))

    def upload_common(self, geo: np.ndarray, scene, rnd: np.ndarray | None = None):
        rnd = random_field(self.spec.random_field_size) if rnd is None else rnd
        geo = np.ascontiguousarray(geo, np.uint8)
        subs = np.ascontiguousarray(scene.substances, np.float32)
        pict = np.ascontiguousarray(scene.pict_data, np.uint8)
        uses = np.ascontiguousarray(scene.pict_uses, np.uint8)
        rnd = np.ascontiguousarray(rnd, np.float32)
        self._check(self.L.gudni_upload_common(
            self.ctx, _ptr(geo), geo.size, _ptr(subs), subs.shape[0], _ptr(pict), pict.size,
            _ptr(uses) if uses.size else None, uses.shape[0], _ptr(rnd), rnd.size))
        grads = _gradients(scene)
        self._check(self.L.gudni_upload_gradients(self.ctx, _ptr(grads) if len(grads) else None,
                                                   len(grads)))

    def raster_job(self, shapes: np.ndarray, tiles: np.ndarray, background, width, height,
                   frame=0, job_index=0, out: np.ndarray | None = None) -> np.ndarray:
        out = np.zeros(width * height, np.uint32) if out is None else out
        bg = np.ascontiguousarray(background, np.float32)
        shapes = np.ascontiguousarray(shapes, np.uint8)
        tiles = np.ascontiguousarray(tiles, np.uint8)
        self._check(self.L.gudni_raster_job(self.ctx, _ptr(shapes) if shapes.size else None,
                                            shapes.size // 16, _ptr(tiles) if tiles.size else None,
                                            tiles.size // 32, _ptr(bg), width, height, frame,
                                            job_index, _ptr(out)))
        return out

    def queue_jobs(self, jobs: RasterJobs, scene, out: np.ndarray | None = None) -> np.ndarray:
        """queueRasterJobs: every job through the Level-1 entry point, numbered
        like the reference's reversed job list."""
        out = np.zeros(scene.width * scene.height, np.uint32) if out is None else out
        for j in range(jobs.n_jobs):
            tiles, shapes = jobs.job(j)
            self.raster_job(shapes, tiles, scene.background, scene.width, scene.height,
                            0, jobs.n_jobs - 1 - j, out)
        return out

    def dump_job(self, shapes, tiles, width, height, job_index=0, cap=64):
        n_tiles = tiles.size // 32
        slots = n_tiles * self.spec.threads_per_tile
        d = {"pre_thr": np.zeros((slots, cap, 4), np.float32), "pre_hdr": np.zeros((slots, cap), np.uint32),
             "thr": np.zeros((slots, cap, 4), np.float32), "hdr": np.zeros((slots, cap), np.uint32),
             "qlen": np.full(slots, -1, np.int32), "shape_bits": np.zeros(slots, np.uint32),
             "stack": np.zeros((slots, 2), np.uint64),
             "shape_indices": np.zeros((slots, self.spec.max_shapes), np.int32)}
        dump = _lib.Dump(cap, *[d[k].ctypes.data for k in
                                ("pre_thr", "pre_hdr", "thr", "hdr", "qlen", "shape_bits", "stack",
                                 "shape_indices")])
        shapes = np.ascontiguousarray(shapes, np.uint8)
        tiles = np.ascontiguousarray(tiles, np.uint8)
        self._check(self.L.gudni_dump_job(self.ctx, _ptr(shapes) if shapes.size else None,
                                          shapes.size // 16, _ptr(tiles), n_tiles, width, height,
                                          job_index, C.byref(dump)))
        return d

    # Level 2 ---------------------------------------------------------------
    def frame_upload(self, jobs: RasterJobs, scene, frame=0):
        bg = np.ascontiguousarray(scene.background, np.float32)
        self._check(self.L.gudni_frame_upload(self.ctx, jobs.handle, scene.width, scene.height,
                                              _ptr(bg), frame))

    def frame_render(self, out_dev_ptr: int, out_f32_dev_ptr: int | None = None):
        self._check(self.L.gudni_frame_render(self.ctx, C.c_void_p(out_dev_ptr),
                                              C.c_void_p(out_f32_dev_ptr) if out_f32_dev_ptr else None))

    def frame_render_static(self, out_dev_ptr: int, out_f32_dev_ptr: int | None = None):
        """gudni_frame_render_static: the uploaded frame rendered with the
        library's static scratch (no device allocation)."""
        self._check(self.L.gudni_frame_render_static(self.ctx, C.c_void_p(out_dev_ptr),
                                                     C.c_void_p(out_f32_dev_ptr) if out_f32_dev_ptr else None))

    def frame_check(self):
        self._check(self.L.gudni_frame_check(self.ctx))

    # Level 2 with stage (a) on the device ------------------------------------
    def _scene_desc(self, geometry: Geometry, scene, rnd):
        rnd = random_field(self.spec.random_field_size) if rnd is None else rnd
        arrs = [np.ascontiguousarray(geometry.geo, np.uint8),
                np.ascontiguousarray(geometry.shapes, np.uint8),
                np.ascontiguousarray(geometry.bbox, np.float32),
                np.ascontiguousarray(scene.substances, np.float32),
                np.ascontiguousarray(scene.pict_data, np.uint8),
                np.ascontiguousarray(scene.pict_uses, np.uint8),
                np.ascontiguousarray(rnd, np.float32), _gradients(scene)]
        geo, shapes, bbox, subs, pict, uses, rnd, grads = arrs
        d = _lib.SceneDesc(_ptr(geo) if geo.size else None, geo.size,
                           _ptr(shapes) if shapes.size else None, shapes.size // 16,
                           _ptr(bbox) if bbox.size else None, _ptr(subs) if subs.size else None,
                           subs.shape[0], _ptr(pict) if pict.size else None, pict.size,
                           _ptr(uses) if uses.size else None, uses.shape[0], _ptr(rnd), rnd.size,
                           scene.width, scene.height,
                           (C.c_float * 4)(*[float(x) for x in scene.background]),
                           _ptr(grads) if len(grads) else None, len(grads))
        return d, arrs

    def scene_upload(self, geometry: Geometry, scene, rnd: np.ndarray | None = None):
        """Scene buffers for device binning: strand heap, shape records and boxes."""
        d, arrs = self._scene_desc(geometry, scene, rnd)
        self._check(self.L.gudni_scene_upload(self.ctx, C.byref(d)))

    def scene_upload_outlines(self, scene, rnd: np.ndarray | None = None):
        """Level 2 from the onShape sequence: outlines go up and the device
        serialises them (strand heap, shape records, boxes), ready for
        scene_bin."""
        rnd = random_field(self.spec.random_field_size) if rnd is None else rnd
        arrs = [np.ascontiguousarray(scene.tags, np.uint64),
                np.ascontiguousarray(scene.shape_outline_offsets, np.int32),
                np.ascontiguousarray(scene.outline_pair_offsets, np.int32),
                np.ascontiguousarray(scene.pairs, np.float32),
                np.ascontiguousarray(scene.substances, np.float32),
                np.ascontiguousarray(scene.pict_data, np.uint8),
                np.ascontiguousarray(scene.pict_uses, np.uint8),
                np.ascontiguousarray(rnd, np.float32), _gradients(scene)]
        tags, soff, ooff, pairs, subs, pict, uses, rnd, grads = arrs
        o = _lib.Outlines(scene.n_shapes, tags.ctypes.data, soff.ctypes.data, ooff.ctypes.data,
                          pairs.ctypes.data if pairs.size else None)
        d = _lib.SceneDesc(None, 0, None, 0, None, _ptr(subs) if subs.size else None, subs.shape[0],
                           _ptr(pict) if pict.size else None, pict.size,
                           _ptr(uses) if uses.size else None, uses.shape[0], _ptr(rnd), rnd.size,
                           scene.width, scene.height,
                           (C.c_float * 4)(*[float(x) for x in scene.background]),
                           _ptr(grads) if len(grads) else None, len(grads))
        self._check(self.L.gudni_scene_upload_outlines(self.ctx, C.byref(o), C.byref(d)))

    def scene_geometry(self):
        """The uploaded scene's strand heap, 16-byte shape records, boxes and
        strand extents as built on the device."""
        sz = np.zeros(2, np.int64)
        self._check(self.L.gudni_scene_geometry_sizes(self.ctx, _ptr(sz)))
        n, nb = int(sz[0]), int(sz[1])
        out = {"geo": np.zeros(nb, np.uint8), "shapes": np.zeros(n * 16, np.uint8),
               "bbox": np.zeros((n, 4), np.float32), "bounds": np.zeros((n, 4), np.float32)}
        self._check(self.L.gudni_scene_geometry_read(self.ctx, _ptr(out["geo"]), _ptr(out["shapes"]),
                                                     _ptr(out["bbox"]), _ptr(out["bounds"])))
        return out

    def last_serialize_ms(self) -> float:
        return float(self.L.gudni_last_serialize_ms(self.ctx))

    def scene_bin(self):
        """Stage (a) on the device; the frame is then ready for frame_render."""
        self._check(self.L.gudni_scene_bin(self.ctx))

    def scene_counts(self):
        c = np.zeros(3, np.int32)
        self._check(self.L.gudni_scene_counts(self.ctx, _ptr(c)))
        return int(c[0]), int(c[1]), int(c[2])

    def scene_read(self):
        """The device-built tiles (shapeStart absolute), reversed job index per
        tile, shape references, their extents and the work list."""
        nt, nr, nw = self.scene_counts()
        out = {"tiles": np.zeros(nt * 32, np.uint8), "tile_job": np.zeros(nt, np.int32),
               "refs": np.zeros(nr * 16, np.uint8), "ref_bounds": np.zeros((nr, 4), np.float32),
               "work": np.zeros(nw, np.uint32)}
        self._check(self.L.gudni_scene_read(
            self.ctx, *[_ptr(out[k]) if out[k].size else None
                        for k in ("tiles", "tile_job", "refs", "ref_bounds", "work")]))
        return out

    def last_bin_ms(self) -> float:
        return float(self.L.gudni_last_bin_ms(self.ctx))

    def raster_scene(self, geometry: Geometry, scene, rnd: np.ndarray | None = None) -> np.ndarray:
        """Upload, bin on the device, render, read back: W*H BGRA words."""
        d, arrs = self._scene_desc(geometry, scene, rnd)
        out = np.zeros(scene.width * scene.height, np.uint32)
        self._check(self.L.gudni_raster_scene(self.ctx, C.byref(d), _ptr(out)))
        return out

    def set_blocks_per_cu(self, n: int):
        """Fast-pass blocks per CU (0 = as many as fit)."""
        self._check(self.L.gudni_set_blocks_per_cu(self.ctx, int(n)))

    def set_output_format(self, fmt: int):
        """FORMAT_BGRA8 (the reference's pixel word, default) or FORMAT_RGBA8."""
        self._check(self.L.gudni_set_output_format(self.ctx, int(fmt)))

    def enable_timing(self, on=True):
        self._check(self.L.gudni_enable_timing(self.ctx, 1 if on else 0))

    def last_timings(self):
        ms = np.zeros(2, np.float32)
        self._check(self.L.gudni_last_timings(self.ctx, _ptr(ms), 2))
        return ms

    def last_stage_cycles(self, n: int = 4):
        """Shader cycles summed over waves.  After a timed render: (generate,
        sort, render, units).  After a timed adjoint run and its result: the
        forward's (generate, sort, sweep), the backward's (generate, sort,
        sweep), the point pass, units."""
        c = np.zeros(n, np.uint64)
        self._check(self.L.gudni_last_stage_cycles(self.ctx, _ptr(c), n))
        return c


def prepare(scene, spec: RasterSpec | None = None):
    """Host stages for a scene: (spec, Geometry, RasterJobs)."""
    spec = spec or default_spec()
    geo = Geometry(scene)
    jobs = RasterJobs(spec, geo, scene.width, scene.height)
    return spec, geo, jobs


class Batch:
    """Many scenes resident on the device, rendered in one persistent launch
    (optionally replayed from a captured hipGraph)."""

    def __init__(self, rasterizer: Rasterizer, items, rnd: np.ndarray | None = None):
        """items: list of (scene, Geometry, RasterJobs)."""
        self._setup(rasterizer, [s for s, _, _ in items], rnd)
        arr = self._scene_array(items)
        self.r._check(self.r.L.gudni_batch_upload(self.r.ctx, len(items), arr, _ptr(self._rnd),
                                                  self._rnd.size, C.byref(self.handle)))

    @classmethod
    def from_outlines(cls, rasterizer: Rasterizer, scenes, tile_sizes=None, radix_bits=None,
                      rnd: np.ndarray | None = None):
        """The batch built on the device from each scene's outlines: serialised
        and binned per scene with its own tile size and sort radix (lists, or
        None for the spec's)."""
        self = cls.__new__(cls)
        self._setup(rasterizer, scenes, rnd)
        arr = self._scene_array([(s, None, None) for s in scenes])
        outl = (_lib.Outlines * len(scenes))()
        for k, sc in enumerate(scenes):
            arrs = [np.ascontiguousarray(sc.tags, np.uint64),
                    np.ascontiguousarray(sc.shape_outline_offsets, np.int32),
                    np.ascontiguousarray(sc.outline_pair_offsets, np.int32),
                    np.ascontiguousarray(sc.pairs, np.float32)]
            self._keep.append(arrs)
            tags, soff, ooff, pairs = arrs
            outl[k] = _lib.Outlines(sc.n_shapes, tags.ctypes.data, soff.ctypes.data, ooff.ctypes.data,
                                    pairs.ctypes.data if pairs.size else None)
        ts = None if tile_sizes is None else np.ascontiguousarray(tile_sizes, np.int32)
        rb = None if radix_bits is None else np.ascontiguousarray(radix_bits, np.int32)
        for a in (ts, rb):
            if a is not None and a.shape != (len(scenes),):
                raise ValueError("one tile size and one radix per scene")
        self._keep += [ts, rb]
        self.r._check(self.r.L.gudni_batch_upload_outlines(self.r.ctx, len(scenes), outl, arr, _ptr(ts), _ptr(rb),
                                                           _ptr(self._rnd), self._rnd.size, C.byref(self.handle)))
        return self

    def _setup(self, rasterizer, scenes, rnd):
        self.r = rasterizer
        self._rnd = random_field(rasterizer.spec.random_field_size) if rnd is None else rnd
        self._keep = [self._rnd]
        self.sizes = [(s.width, s.height) for s in scenes]
        self.handle = C.c_void_p()

    def _scene_array(self, items):
        arr = (_lib.BatchScene * len(items))()
        for k, (scene, geo, jobs) in enumerate(items):
            subs = np.ascontiguousarray(scene.substances, np.float32)
            pict = np.ascontiguousarray(scene.pict_data, np.uint8)
            uses = np.ascontiguousarray(scene.pict_uses, np.uint8)
            grads = _gradients(scene)
            self._keep += [geo, jobs, subs, pict, uses, grads]
            a = arr[k]
            if geo is not None:
                a.geo = geo.geo.ctypes.data if geo.geo.size else None
                a.geo_bytes = geo.geo.size
                a.jobs = jobs.handle
            a.substances = subs.ctypes.data
            a.n_substances = subs.shape[0]
            a.pict_data = pict.ctypes.data if pict.size else None
            a.pict_bytes = pict.size
            a.pict_uses = uses.ctypes.data if uses.size else None
            a.n_pict_uses = uses.shape[0]
            a.gradients = grads.ctypes.data if len(grads) else None
            a.n_gradients = len(grads)
            a.width, a.height = scene.width, scene.height
            for c in range(4):
                a.background[c] = float(scene.background[c])
        return arr

    def render(self, use_graph: bool = False):
        self.r._check(self.r.L.gudni_batch_render(self.r.ctx, self.handle, 1 if use_graph else 0))

    def render_scene(self, i: int):
        self.r._check(self.r.L.gudni_batch_render_scene(self.r.ctx, self.handle, i))

    def check(self):
        lat = np.zeros(len(self.sizes), np.float32)
        self.r._check(self.r.L.gudni_batch_check(self.r.ctx, self.handle, _ptr(lat)))
        return lat

    def frame(self, i: int) -> np.ndarray:
        w, h = self.sizes[i]
        out = np.zeros(w * h, np.uint32)
        self.r._check(self.r.L.gudni_batch_frame(self.r.ctx, self.handle, i, _ptr(out)))
        return out.reshape(h, w)

    def close(self):
        if self.handle:
            self.r.L.gudni_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """Gradients of the L2 image loss for one or more scenes (one launch runs
    them all).  items: (scene, target) pairs, target an (H, W, 4) float32
    image; the scene's outline pairs and substance colours are the
    parameters."""

    def __init__(self, rasterizer: Rasterizer, items):
        self.r = rasterizer
        self.L = rasterizer.L
        self.scenes = [sc for sc, _ in items]
        self._keep = []
        descs = (_lib.AdjointScene * len(items))()
        for i, (sc, target) in enumerate(items):
            arrs = [np.ascontiguousarray(sc.tags, np.uint64),
                    np.ascontiguousarray(sc.shape_outline_offsets, np.int32),
                    np.ascontiguousarray(sc.outline_pair_offsets, np.int32),
                    np.ascontiguousarray(sc.pairs, np.float32),
                    np.ascontiguousarray(sc.substances, np.float32),
                    np.ascontiguousarray(sc.pict_data, np.uint8),
                    np.ascontiguousarray(sc.pict_uses, np.uint8),
                    np.ascontiguousarray(target, np.float32), _gradients(sc)]
            assert arrs[7].shape == (sc.height, sc.width, 4)
            self._keep.append(arrs)
            tags, soff, ooff, pairs, subs, pict, uses, tgt, grads = arrs
            o = _lib.Outlines(sc.n_shapes, *[a.ctypes.data for a in (tags, soff, ooff, pairs)])
            descs[i] = _lib.AdjointScene(o, _ptr(subs) if subs.size else None, subs.shape[0],
                                         _ptr(pict) if pict.size else None, pict.size,
                                         _ptr(uses) if uses.size else None, uses.size // 24,
                                         _ptr(tgt), sc.width, sc.height,
                                         (C.c_float * 4)(*[float(x) for x in sc.background]),
                                         _ptr(grads) if len(grads) else None, len(grads))
        self.handle = C.c_void_p()
        self.r._check(self.L.gudni_adjoint_upload(self.r.ctx, len(items), descs, C.byref(self.handle)))

    def update(self, i, pairs=None, rgba=None):
        """New parameters for scene i (re-serialised and re-binned)."""
        p = None if pairs is None else np.ascontiguousarray(pairs, np.float32)
        c = None if rgba is None else np.ascontiguousarray(rgba, np.float32)
        self.r._check(self.L.gudni_adjoint_update(self.r.ctx, self.handle, i, _ptr(p), _ptr(c)))

    def update_device(self, i, pairs_ptr=None, rgba_ptr=None):
        """New parameters for scene i from device memory (e.g. a torch
        tensor's data_ptr()); re-serialised and re-binned on the device."""
        self.r._check(self.L.gudni_adjoint_update_device(self.r.ctx, self.handle, i,
                                                         C.c_void_p(pairs_ptr) if pairs_ptr else None,
                                                         C.c_void_p(rgba_ptr) if rgba_ptr else None))

    def result_device(self, i, loss_ptr=None, d_pairs_ptr=None, d_rgba_ptr=None, image_ptr=None):
        """Stream-ordered copies of scene i's loss (f64), gradients and image
        into device memory; errors surface at the next result() or check()."""
        p = [C.c_void_p(x) if x else None for x in (loss_ptr, d_pairs_ptr, d_rgba_ptr, image_ptr)]
        self.r._check(self.L.gudni_adjoint_result_device(self.r.ctx, self.handle, i, *p))

    def check(self):
        self.r._check(self.L.gudni_adjoint_check(self.r.ctx, self.handle))

    def run(self):
        self.r._check(self.L.gudni_adjoint_run(self.r.ctx, self.handle))

    def run_scene(self, i):
        self.r._check(self.L.gudni_adjoint_run_scene(self.r.ctx, self.handle, i))

    def result(self, i, image=False):
        """(loss, d_pairs (n_pairs, 4), d_rgba (n_substances, 4), image or None)."""
        sc = self.scenes[i]
        loss = C.c_double()
        dp = np.zeros((sc.pairs.size // 4, 4), np.float32)
        dc = np.zeros((len(sc.substances), 4), np.float32)
        img = np.zeros((sc.height, sc.width, 4), np.float32) if image else None
        self.r._check(self.L.gudni_adjoint_result(self.r.ctx, self.handle, i, C.byref(loss),
                                                  _ptr(dp) if dp.size else None,
                                                  _ptr(dc) if dc.size else None, _ptr(img)))
        return loss.value, dp, dc, img

    def close(self):
        if self.handle:
            self.L.gudni_adjoint_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
