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


def test_batch_sequential_scenes(batch_and_refs):
    r, b, scenes, refs = batch_and_refs
    say("sequential per-scene launches")
    for i in range(len(scenes)):
        b.render_scene(i)
    b.check()
    check_all(b, scenes, refs, "sequential")


@pytest.mark.parametrize("tile", [64, 128])
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
