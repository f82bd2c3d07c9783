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


def test_work_order_longest_first():
    """gudni_order_work: a permutation of the work list; the units within a
    factor 8 of the costliest (rows per lane x ((shapes + 1) / width)^2) open
    the eight claimer ranges, longest first, a tile's quarters together; all
    other units keep their original order."""
    rng = np.random.default_rng(4)
    n_tiles = 300
    tiles = np.zeros((n_tiles, 8), np.int32)
    hd = rng.integers(3, 9, n_tiles)
    tiles[:, 2] = 1 << hd                                   # box: 0 0 w h
    tiles[:, 3] = 1 << rng.integers(3, 9, n_tiles)
    tiles[:, 4] = (hd | (hd << 16)).astype(np.int32)        # hdepth, vdepth
    tiles[:, 7] = rng.integers(0, 300, n_tiles)             # shape count
    work = np.array([(t << 2) | q for t in range(n_tiles) for q in range(4) if q < max(1, (1 << hd[t]) // 64)],
                    np.uint32)
    out = np.zeros_like(work)
    L = _lib.load()
    assert L.gudni_order_work(tiles.ctypes.data, n_tiles, None, work.ctypes.data, work.size, 256, 8,
                              out.ctypes.data) == 0
    assert sorted(out.tolist()) == sorted(work.tolist())

    def cost(units):
        t = units >> 2
        rows = np.maximum(1, tiles[t, 3] >> np.maximum(0, 8 - hd[t]))
        return rows * ((tiles[t, 7] + 1.0) / (1 << hd[t])) ** 2

    c = cost(out)
    heavy = c * 8 >= c.max()
    n = work.size
    bounds = [n * p // 8 for p in range(9)]
    for p in range(8):
        h = heavy[bounds[p]:bounds[p + 1]]
        k = int(h.sum())
        assert h[:k].all() and not h[k:].any(), "long units open the range"
        assert np.all(np.diff(c[bounds[p]:bounds[p] + k]) <= 0), "longest first"
    assert heavy[bounds[:8]].all()
    # the others in their original order
    pos = {int(u): i for i, u in enumerate(work)}
    rest = [pos[int(u)] for u, hv in zip(out, heavy) if not hv]
    assert rest == sorted(rest)
    # the quarters of a tile stay adjacent, in order
    t = out >> 2
    for i in range(1, n):
        if t[i] == t[i - 1]:
            assert (out[i] & 3) == (out[i - 1] & 3) + 1
    bad = np.zeros(4, np.uint32)
    assert L.gudni_order_work(tiles.ctypes.data, 1, None, work.ctypes.data, 4, 256, 8, bad.ctypes.data) != 0
    deep = tiles.copy()
    deep[:, 4] = 40                                         # hdepth beyond any tile width
    assert L.gudni_order_work(deep.ctypes.data, n_tiles, None, work.ctypes.data, work.size, 256, 8,
                              out.ctypes.data) != 0


def test_work_order_cheapest_scene_first():
    """gudni_order_work with the tiles' scenes: the long units open the
    ranges as without them; after them each range holds whole runs of scenes
    in order of their summed cost, cheapest first, each scene's units in
    their original order, and the scenes are dealt round robin."""
    rng = np.random.default_rng(9)
    n_scenes = 40
    tiles, tile_scene = [], []
    for sc in range(n_scenes):
        k = int(rng.integers(1, 30))
        for _ in range(k):
            hd = int(rng.integers(3, 7))
            tiles.append([0, 0, 1 << hd, 1 << hd, hd | (hd << 16), 0, 0, int(rng.integers(0, 120))])
            tile_scene.append(sc)
    tiles = np.array(tiles, np.int32)
    tile_scene = np.array(tile_scene, np.int32)
    hd = tiles[:, 4] & 0xFFFF
    n_tiles = len(tiles)
    work = np.array([(t << 2) | q for t in range(n_tiles) for q in range(4) if q < max(1, (1 << hd[t]) // 64)],
                    np.uint32)
    out = np.zeros_like(work)
    L = _lib.load()
    assert L.gudni_order_work(tiles.ctypes.data, n_tiles, tile_scene.ctypes.data, work.ctypes.data, work.size,
                              256, 8, out.ctypes.data) == 0
    assert sorted(out.tolist()) == sorted(work.tolist())
    t = (work >> 2).astype(np.int64)
    rows = np.maximum(1, tiles[t, 3] >> np.maximum(0, 8 - hd[t]))
    cost = dict(zip(work.tolist(), (rows * ((tiles[t, 7] + 1.0) / (1 << hd[t])) ** 2).tolist()))
    top = max(cost.values())
    heavy = {u for u, c in cost.items() if c * 8 >= top}
    scene_cost = {}
    for u, c in cost.items():
        if u not in heavy:
            scene_cost[int(tile_scene[u >> 2])] = scene_cost.get(int(tile_scene[u >> 2]), 0.0) + c
    pos = {int(u): i for i, u in enumerate(work)}
    n = work.size
    bounds = [n * p // 8 for p in range(9)]
    first_scene = []
    for p in range(8):
        rng_units = [int(u) for u in out[bounds[p]:bounds[p + 1]]]
        k = sum(u in heavy for u in rng_units)
        assert all(u in heavy for u in rng_units[:k]), "long units open the range"
        rest = rng_units[k:]
        costs = [scene_cost[int(tile_scene[u >> 2])] for u in rest]
        assert costs == sorted(costs), "cheapest scenes first"
        for sc in {int(tile_scene[u >> 2]) for u in rest}:
            mine = [pos[u] for u in rest if tile_scene[u >> 2] == sc]
            assert mine == sorted(mine), "a scene's units keep their order"
        if rest:
            first_scene.append(int(tile_scene[rest[0] >> 2]))
    # round robin: the eight cheapest scenes open the eight ranges' scene runs
    cheapest = sorted(scene_cost, key=lambda sc: (scene_cost[sc], sc))[:8]
    assert sorted(first_scene) == sorted(cheapest)


def test_build_fingerprint_tracks_sources(tmp_path, monkeypatch):
    """A library whose stamp does not match the sources and flags of the
    tree is stale, so tests and loaders rebuild it instead of running it."""
    from gudni_amd import build as pbuild
    from oracle import build as obuild
    assert not pbuild.stale() and not obuild.stale()
    fp = pbuild.fingerprint()
    assert fp == pbuild.fingerprint()
    src = tmp_path / "extra.hip"
    src.write_text("// a\n")
    monkeypatch.setattr(pbuild, "SOURCES", pbuild.SOURCES + [str(src)])
    fa = pbuild.fingerprint()
    assert fa != fp and pbuild.stale()
    src.write_text("// b\n")
    assert pbuild.fingerprint() != fa
    monkeypatch.setattr(pbuild, "SOURCES", pbuild.SOURCES[:-1])
    monkeypatch.setattr(pbuild, "ARCH", "gfx942")
    assert pbuild.fingerprint() != fp and pbuild.stale()


def test_build_fingerprint_moves_with_the_tree(tmp_path):
    """The same sources and flags in a copy of the tree at another path have
    the same fingerprint, so a tree built here is not rebuilt where it runs."""
    import shutil
    import subprocess
    import sys
    from gudni_amd import build as pbuild
    from oracle import build as obuild
    root = pbuild.ROOT
    copy = tmp_path / "elsewhere"
    for d in ("gudni_amd", "include", "oracle"):
        shutil.copytree(os.path.join(root, d), copy / d,
                        ignore=shutil.ignore_patterns("*.so", "*.stamp", "__pycache__"))
    code = ("from gudni_amd import build as B; from oracle import build as O; "
            "print(B.fingerprint()); print(O.fingerprint())")
    out = subprocess.run([sys.executable, "-c", code], cwd=copy, capture_output=True, text=True, check=True)
    assert out.stdout.split() == [pbuild.fingerprint(), obuild.fingerprint()]


def test_work_units_cover_every_thread_once():
    """The four work units of a tile run each of its 256 threads (reference
    column indices) exactly once, in blocks as square as the tile allows."""
    from tests.helpers import unit_columns
    lanes = np.arange(64)
    for hd in range(2, 9):
        cols = np.concatenate([unit_columns(hd, 8, q, lanes) for q in range(4)])
        assert np.array_equal(np.sort(cols), np.arange(256)), hd
        for q in range(4):
            c = unit_columns(hd, 8, q, lanes)
            nx, ns = len(np.unique(c & ((1 << hd) - 1))), len(np.unique(c >> hd))
            assert nx * ns == 64
            if 3 <= hd <= 5:
                assert nx == 8 and ns == 8, (hd, nx, ns)
