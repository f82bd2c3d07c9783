"""Stage (a) on the device: the tile tree, per-tile shape references, job
fields and work list built by the binning kernels equal the host build
(which the oracle's literal insert/split tile tree pins) byte for byte, and the
frame rendered from them equals the oracle's."""
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from tests import helpers as H
from tests.test_gpu_parity import comb_scene, compound_scene, textured_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-binning] {msg}", flush=True)
    sys.stdout.flush()


def empty_scene():
    return S.build_scene(300, 200, [], [[1, 0, 0, 1]], background=(0.2, 0.3, 0.4, 1.0), name="empty")


SCENES = {
    "tinySquare": H.tiny_square,
    "compound": compound_scene,
    "textured": textured_scene,
    "comb": comb_scene,
    "empty": empty_scene,
    "uniform_512": lambda: S.uniform_512(),
    "batch_uneven": lambda: S.batch_scene(640, 5000, 3),
    "blobs_1024": lambda: S.blobs_scaled(1024, 12500, 7),
}


def compare_binning(r, scene, spec):
    _, geo, jobs = G.prepare(scene, spec)
    r.scene_upload(geo, scene)
    r.scene_bin()
    got = r.scene_read()
    tiles, tile_job, refs, work = H.host_frame_layout(spec, jobs, scene.width, scene.height)
    nt, nr, nw = r.scene_counts()
    assert (nt, nr) == (jobs.n_tiles, jobs.n_refs), f"{scene.name}: counts {(nt, nr)} vs host {(jobs.n_tiles, jobs.n_refs)}"
    gt = got["tiles"].view(H.TILE_DT)
    for f in H.TILE_DT.names:
        bad = np.argwhere(np.atleast_2d((gt[f] != tiles[f]).reshape(len(tiles), -1)).any(axis=1))
        assert bad.size == 0, f"{scene.name}: tile field {f} differs first at tile {int(bad[0][0])}"
    np.testing.assert_array_equal(got["tile_job"], tile_job, err_msg="tile_job")
    np.testing.assert_array_equal(got["refs"], refs, err_msg="shape references")
    np.testing.assert_array_equal(got["work"], work, err_msg="work list")
    return jobs


@pytest.mark.parametrize("name", list(SCENES))
def test_device_binning_equals_host(rasterizer, name):
    scene = SCENES[name]()
    say(f"{name} {scene.width}x{scene.height} {scene.n_shapes} shapes")
    compare_binning(rasterizer, scene, rasterizer.spec)


@pytest.mark.parametrize("tile", [32, 64, 128])
def test_device_binning_other_tile_sizes(tile):
    spec = G.default_spec(max_tile_size=tile)
    r = G.Rasterizer(spec)
    try:
        for scene in (S.batch_scene(320, 800, 5), S.uniform_512(120, 160, 6)):
            say(f"tile {tile}: {scene.name}")
            compare_binning(r, scene, spec)
    finally:
        r.close()


def test_device_binning_full_blobs_4096(rasterizer):
    scene = S.blobs_4096()
    say("blobs 4096^2 device binning")
    compare_binning(rasterizer, scene, rasterizer.spec)


def test_raster_scene_matches_oracle(rasterizer):
    """gudni_raster_scene (upload + device binning + render + read back)."""
    for scene in (textured_scene(), S.uniform_512(200, 128, 3), S.batch_scene(384, 2000, 2)):
        say(f"raster_scene {scene.name}")
        _, geo, _ = G.prepare(scene)
        out = rasterizer.raster_scene(geo, scene).reshape(scene.height, scene.width)
        ref = H.oracle_render(scene, want_f64=False)
        bad = H.first_mismatch(out, ref["out"])
        assert bad is None, f"{scene.name}: first mismatch {bad}"
