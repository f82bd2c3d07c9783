"""The C ABI: the library loads without a GPU, exports every function the
header declares, and its host-only entry points reject bad arguments."""
import ctypes as C
import os
import re

import numpy as np

import gudni_amd as G
from gudni_amd import _lib
from gudni_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gudni_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(gudni_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_signatures_cover_the_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_struct_sizes_match_the_header():
    assert C.sizeof(_lib.Spec) == 44
    assert C.sizeof(_lib.BatchScene) == 112
    assert C.sizeof(_lib.SceneDesc) == 144
    assert C.sizeof(_lib.AdjointScene) == 136
    assert C.sizeof(_lib.Outlines) == 40
    assert C.sizeof(_lib.Dump) == 72


def test_default_spec_is_the_pinned_spec():
    s = G.default_spec()
    assert (s.max_tile_size, s.threads_per_tile, s.max_tiles_per_call, s.max_thresholds,
            s.max_strands_per_tile, s.max_shapes, s.random_field_size) == (256, 256, 256, 4096, 4094, 127, 4096)
    assert s.stochastic_factor == 0.0 and s.coverage_mode == 0 and s.texture_filter == 0
    assert s.sort_radix_bits == 4


def test_host_entry_points_reject_bad_arguments():
    L = _lib.load()
    h = C.c_void_p()
    assert L.gudni_geometry_build(None, 10, 10, C.byref(h)) == -1
    scene = S.uniform_512(5, 32, 1)
    geo = G.Geometry(scene)
    spec = G.default_spec()
    assert L.gudni_jobs_build(C.byref(spec), geo.handle, 0, 10, C.byref(h)) == -1
    bad = G.default_spec(max_tile_size=100)
    assert L.gudni_jobs_build(C.byref(bad), geo.handle, 32, 32, C.byref(h)) == -1
    assert b"bad spec" in L.gudni_last_error(None)


def test_geometry_culls_offcanvas_and_empty_shapes():
    shapes = [(S.solid_tag(0), [S.rectangle_pairs(1, 1, 5, 5)]),
              (S.solid_tag(1), [S.rectangle_pairs(40, 1, 50, 5)]),      # right of canvas
              (S.solid_tag(2), [S.rectangle_pairs(-9, -9, 0, 0)]),      # touches at 0: culled
              (S.solid_tag(3), [np.full((1, 4), 5.0, np.float32)])]     # one pair: no strands
    scene = S.build_scene(32, 32, shapes, np.ones((4, 4), np.float32))
    geo = G.Geometry(scene)
    assert list(geo.source_index) == [0, 3]
    assert list(geo.strand_count) == [4, 0]


def test_jobs_reject_tiles_wider_than_the_threads():
    L = _lib.load()
    h = C.c_void_p()
    geo = G.Geometry(S.uniform_512(5, 32, 1))
    wide = G.default_spec(max_tile_size=512)
    assert L.gudni_jobs_build(C.byref(wide), geo.handle, 32, 32, C.byref(h)) == -1
    assert b"bad spec" in L.gudni_last_error(None)
