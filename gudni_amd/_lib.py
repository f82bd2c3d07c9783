"""ctypes declarations of libgudni_hip's C ABI (include/gudni_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgudni_hip.so")


class Spec(C.Structure):
    _fields_ = [("max_tile_size", C.c_int32), ("threads_per_tile", C.c_int32),
                ("max_tiles_per_call", C.c_int32), ("max_thresholds", C.c_int32),
                ("max_strands_per_tile", C.c_int32), ("max_shapes", C.c_int32),
                ("random_field_size", C.c_int32), ("stochastic_factor", C.c_float),
                ("coverage_mode", C.c_int32), ("texture_filter", C.c_int32),
                ("sort_radix_bits", C.c_int32)]


class Outlines(C.Structure):
    _fields_ = [("n_shapes", C.c_int32), ("tags", C.c_void_p),
                ("shape_outline_offsets", C.c_void_p), ("outline_pair_offsets", C.c_void_p),
                ("pairs", C.c_void_p)]


class GeometryView(C.Structure):
    _fields_ = [("geo", C.POINTER(C.c_uint8)), ("geo_bytes", C.c_size_t),
                ("shapes", C.POINTER(C.c_uint8)), ("bbox", C.POINTER(C.c_float)),
                ("strand_count", C.POINTER(C.c_uint32)), ("source_index", C.POINTER(C.c_int32)),
                ("n_shapes", C.c_int32)]


class JobsView(C.Structure):
    _fields_ = [("n_jobs", C.c_int32), ("job_tile_offset", C.POINTER(C.c_int32)),
                ("job_shape_offset", C.POINTER(C.c_int32)), ("tiles", C.POINTER(C.c_uint8)),
                ("shape_refs", C.POINTER(C.c_uint8)), ("n_tiles", C.c_int32),
                ("n_refs", C.c_int32)]


class Dump(C.Structure):
    _fields_ = [("cap", C.c_int32), ("pre_thresholds", C.c_void_p), ("pre_headers", C.c_void_p),
                ("thresholds", C.c_void_p), ("headers", C.c_void_p), ("queue_len", C.c_void_p),
                ("shape_bits", C.c_void_p), ("shape_stack", C.c_void_p),
                ("shape_indices", C.c_void_p)]


class BatchScene(C.Structure):
    _fields_ = [("geo", C.c_void_p), ("geo_bytes", C.c_size_t), ("substances", C.c_void_p),
                ("n_substances", C.c_int32), ("pict_data", C.c_void_p), ("pict_bytes", C.c_size_t),
                ("pict_uses", C.c_void_p), ("n_pict_uses", C.c_int32), ("jobs", C.c_void_p),
                ("width", C.c_int32), ("height", C.c_int32), ("background", C.c_float * 4),
                ("gradients", C.c_void_p), ("n_gradients", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("geo", C.c_void_p), ("geo_bytes", C.c_size_t), ("shapes", C.c_void_p),
                ("n_shapes", C.c_int32), ("bbox", C.c_void_p), ("substances", C.c_void_p),
                ("n_substances", C.c_int32), ("pict_data", C.c_void_p), ("pict_bytes", C.c_size_t),
                ("pict_uses", C.c_void_p), ("n_pict_uses", C.c_int32), ("random_field", C.c_void_p),
                ("n_random", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("background", C.c_float * 4), ("gradients", C.c_void_p), ("n_gradients", C.c_int32)]


class AdjointScene(C.Structure):
    _fields_ = [("outlines", Outlines), ("rgba", C.c_void_p), ("n_substances", C.c_int32),
                ("pict_data", C.c_void_p), ("pict_bytes", C.c_size_t), ("pict_uses", C.c_void_p),
                ("n_pict_uses", C.c_int32), ("target", C.c_void_p), ("width", C.c_int32),
                ("height", C.c_int32), ("background", C.c_float * 4), ("gradients", C.c_void_p),
                ("n_gradients", C.c_int32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "gudni_default_spec": (None, [C.POINTER(Spec)]),
    "gudni_create": (C.c_int, [C.POINTER(Spec), C.c_int, C.POINTER(C.c_void_p)]),
    "gudni_destroy": (None, [C.c_void_p]),
    "gudni_last_error": (C.c_char_p, [C.c_void_p]),
    "gudni_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gudni_upload_common": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_int32,
                                      C.c_void_p, C.c_size_t, C.c_void_p, C.c_int32, C.c_void_p,
                                      C.c_int32]),
    "gudni_upload_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "gudni_raster_job": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                   C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_void_p]),
    "gudni_geometry_build": (C.c_int, [C.POINTER(Outlines), C.c_int32, C.c_int32,
                                       C.POINTER(C.c_void_p)]),
    "gudni_geometry_view_get": (C.c_int, [C.c_void_p, C.POINTER(GeometryView)]),
    "gudni_geometry_free": (None, [C.c_void_p]),
    "gudni_jobs_build": (C.c_int, [C.POINTER(Spec), C.c_void_p, C.c_int32, C.c_int32,
                                   C.POINTER(C.c_void_p)]),
    "gudni_jobs_view_get": (C.c_int, [C.c_void_p, C.POINTER(JobsView)]),
    "gudni_jobs_free": (None, [C.c_void_p]),
    "gudni_frame_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                     C.c_int32]),
    "gudni_frame_render": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gudni_frame_render_static": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gudni_frame_check": (C.c_int, [C.c_void_p]),
    "gudni_enable_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "gudni_set_blocks_per_cu": (C.c_int, [C.c_void_p, C.c_int32]),
    "gudni_set_output_format": (C.c_int, [C.c_void_p, C.c_int32]),
    "gudni_last_timings": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "gudni_last_stage_cycles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "gudni_dump_job": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                 C.c_int32, C.c_int32, C.c_int32, C.POINTER(Dump)]),
    "gudni_batch_upload": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(BatchScene), C.c_void_p,
                                     C.c_int32, C.POINTER(C.c_void_p)]),
    "gudni_batch_upload_outlines": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Outlines), C.POINTER(BatchScene),
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                              C.POINTER(C.c_void_p)]),
    "gudni_batch_render": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "gudni_batch_render_scene": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "gudni_batch_check": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "gudni_batch_frame": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "gudni_batch_free": (None, [C.c_void_p]),
    "gudni_order_work": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_void_p]),
    "gudni_scene_upload": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc)]),
    "gudni_scene_bin": (C.c_int, [C.c_void_p]),
    "gudni_scene_counts": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gudni_scene_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]),
    "gudni_last_bin_ms": (C.c_float, [C.c_void_p]),
    "gudni_adjoint_update_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "gudni_adjoint_result_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p]),
    "gudni_adjoint_check": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gudni_scene_upload_outlines": (C.c_int, [C.c_void_p, C.POINTER(Outlines), C.POINTER(SceneDesc)]),
    "gudni_scene_geometry_sizes": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gudni_scene_geometry_read": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gudni_last_serialize_ms": (C.c_float, [C.c_void_p]),
    "gudni_raster_scene": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.c_void_p]),
    "gudni_adjoint_upload": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(AdjointScene),
                                       C.POINTER(C.c_void_p)]),
    "gudni_adjoint_update": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "gudni_adjoint_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gudni_adjoint_run_scene": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "gudni_adjoint_result": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    "gudni_adjoint_free": (None, [C.c_void_p]),
    "gudni_adjoint": (C.c_int, [C.c_void_p, C.POINTER(AdjointScene), C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p]),
}

_lib = None


def load():
    """Load the in-tree library.  There is no fallback: a missing library is
    an error."""
    global _lib
    if _lib is None:
        # libgudni_hip and torch both need libamdhip64.so.7, and the first
        # copy loaded serves the whole process.  torch ships its own build of
        # that runtime and fails to initialise on a different one, so when
        # torch is installed it is imported first and this library binds to
        # the same runtime torch uses.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        from . import build as _build
        if _build.stale():
            # like the reference, which compiles its kernels at every setup
            # (OpenCL/Setup.hs:124-134), never run a binary that does not
            # match the sources: rebuild (there is no fallback if that fails)
            print("[gudni] libgudni_hip.so is missing or older than its sources: building",
                  file=sys.stderr, flush=True)
            _build.build()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
