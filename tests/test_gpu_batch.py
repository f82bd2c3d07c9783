"""Batched rendering: many scenes of uneven size in one persistent launch,
directly or replayed from a captured hipGraph.  Every scene's frame must
equal the oracle's render of that scene, bit for bit."""
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import helpers as H
from tests.test_gpu_parity import comb_scene, compound_scene, textured_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-batch] {msg}", flush=True)
    sys.stdout.flush()


def batch_scenes():
    return [S.batch_scene(256, 300, 1), S.batch_scene(384, 2000, 2), textured_scene(),
            compound_scene(), comb_scene(), S.uniform_512(150, 200, 9), S.batch_scene(640, 5000, 3),
            H.tiny_square()]


@pytest.fixture(scope="module")
def batch_and_refs():
    import torch
    r = G.Rasterizer()
    r.set_stream(torch.cuda.current_stream().cuda_stream)
    scenes = batch_scenes()
    items = [(sc,) + G.prepare(sc)[1:] for sc in scenes]
    say(f"upload {len(items)} scenes")
    b = G.Batch(r, items)
    refs = [H.oracle_render(sc, want_f64=False)["out"] for sc in scenes]
    yield r, b, scenes, refs
    b.close()
    r.close()


def check_all(b, scenes, refs, what):
    for i, (sc, ref) in enumerate(zip(scenes, refs)):
        got = b.frame(i)
        bad = H.first_mismatch(got, ref)
        assert bad is None, f"{what}: scene {i} ({sc.name}) first mismatch {bad}"


def test_batch_direct(batch_and_refs):
    r, b, scenes, refs = batch_and_refs
    say("direct launch")
    b.render(use_graph=False)
    lat = b.check()
    check_all(b, scenes, refs, "direct")
    assert np.all(lat > 0)


def test_batch_graph_replays(batch_and_refs):
    """Capture once, replay three times with eager work and a synchronisation
    between replays; every replay must produce the oracle's frames."""
    import torch
    r, b, scenes, refs = batch_and_refs
    for k in range(3):
        say(f"graph replay {k}")
        b.render(use_graph=True)
        b.check()
        check_all(b, scenes, refs, f"replay {k}")
        # eager work on the same stream between replays, then clobber one frame
        x = torch.ones(1 << 20, device="cuda")
        torch.cuda.synchronize()
        assert float(x.sum()) == float(1 << 20)
        b.render_scene(0)
        b.check()


def test_batch_graph_recaptured_after_scratch_change(batch_and_refs):
    """A captured graph holds the ring scratch and the fast grid: after
    gudni_set_blocks_per_cu reallocates them, the next replay must be
    re-captured against the new buffers and still give the oracle's frames."""
    r, b, scenes, refs = batch_and_refs
    try:
        for bpc in (2, 6, 0):
            say(f"graph after set_blocks_per_cu({bpc})")
            b.render(use_graph=True)
            b.check()
            r.set_blocks_per_cu(bpc)
            b.render(use_graph=True)
            b.check()
            check_all(b, scenes, refs, f"replay after blocks_per_cu={bpc}")
    finally:
        r.set_blocks_per_cu(0)


def test_batch_sequential_scenes(batch_and_refs):
    r, b, scenes, refs = batch_and_refs
    say("sequential per-scene launches")
    for i in range(len(scenes)):
        b.render_scene(i)
    b.check()
    check_all(b, scenes, refs, "sequential")


@pytest.mark.parametrize("tile", [32, 64, 128])
def test_batch_other_tile_sizes(tile):
    """The autotuner's tile sizes: frames equal the oracle's render under the
    same spec (its own tile tree)."""
    import torch
    spec = G.default_spec(max_tile_size=tile)
    r = G.Rasterizer(spec)
    try:
        r.set_stream(torch.cuda.current_stream().cuda_stream)
        scenes = [S.batch_scene(320, 800, 5), S.uniform_512(120, 160, 6), textured_scene()]
        items = [(sc,) + G.prepare(sc, spec)[1:] for sc in scenes]
        say(f"tile {tile}: upload")
        b = G.Batch(r, items)
        b.render(use_graph=True)
        b.check()
        for i, sc in enumerate(scenes):
            ref = H.oracle_render(sc, spec, want_f64=False)["out"]
            bad = H.first_mismatch(b.frame(i), ref)
            assert bad is None, f"tile {tile}: scene {sc.name} mismatch {bad}"
        b.close()
    finally:
        r.close()


def test_batch_from_outlines_equals_host_batch():
    """The batch serialised and binned on the device, each scene with its own
    tile size and sort radix, renders the frames of the batch built on the
    host with the same per-scene specs, and the oracle's."""
    import torch
    from tests.test_gpu_binning import empty_scene
    from tests.test_gpu_serialize import awkward_scene
    # scenes from empty to 1400^2 with 20,000 shapes
    scenes = batch_scenes() + [empty_scene(), awkward_scene(with_nan=False), S.blobs_scaled(1024, 12500, 7),
                               S.blobs_scaled(1400, 20000, 9)]
    tiles = [256, 64, 128, 256, 64, 128, 256, 32, 128, 64, 256, 64]
    radix = [0, 3, 5, 4, 0, 3, 5, 2, 1, 0, 3, 3]
    assert len(tiles) == len(radix) == len(scenes)
    r = G.Rasterizer()
    try:
        r.set_stream(torch.cuda.current_stream().cuda_stream)
        items = [(sc,) + G.prepare(sc, G.default_spec(max_tile_size=t, sort_radix_bits=q))[1:]
                 for sc, t, q in zip(scenes, tiles, radix)]
        say(f"host-built batch of {len(items)} scenes")
        host = G.Batch(r, items)
        host.render()
        host.check()
        want = [host.frame(i) for i in range(len(scenes))]
        host.close()
        for rep in range(2):
            say(f"device-built batch, upload {rep}")
            b = G.Batch.from_outlines(r, scenes, tiles, radix)
            for use_graph in (False, True, True):
                b.render(use_graph=use_graph)
                b.check()
                for i, sc in enumerate(scenes):
                    got = b.frame(i)
                    assert np.array_equal(got, want[i]), \
                        f"scene {i} ({sc.name}, tile {tiles[i]}) differs from the host-built batch"
            b.close()
        for i in (0, 2, 9):
            spec = G.default_spec(max_tile_size=tiles[i])
            ref = H.oracle_render(scenes[i], spec, want_f64=False)["out"]
            bad = H.first_mismatch(want[i], ref)
            assert bad is None, f"scene {scenes[i].name}: mismatch with the oracle {bad}"
    finally:
        r.close()


def test_batch_from_outlines_rejects_bad_tiles():
    r = G.Rasterizer()
    try:
        scenes = [H.tiny_square(), S.uniform_512(100, 100, 1)]
        for tiles in ([256, 48], [512, 256], [256, 4]):
            with pytest.raises(G.GudniError) as ei:
                G.Batch.from_outlines(r, scenes, tiles, None)
            assert "tile size" in str(ei.value)
        with pytest.raises(G.GudniError):
            G.Batch.from_outlines(r, scenes, None, [0, 6])
        b = G.Batch.from_outlines(r, scenes)
        b.render()
        b.check()
        ref = H.oracle_render(scenes[0], want_f64=False)["out"]
        assert H.first_mismatch(b.frame(0), ref) is None
        b.close()
    finally:
        r.close()
