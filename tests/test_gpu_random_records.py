"""A seeded sweep of tile assignment and sort order under random specs: for
small scenes with random tile size and sort radix, (1) the product's host tile
tree and jobs equal the oracle's literal insert/split procedure byte for byte,
(2) the device-built tiles, references, job numbering and work list equal
the host build, and (3) every column's records after generate and after the
radix pass + sort, its shape bits, initial stack and shape indices equal the
oracle's, on the first and the last job."""
import sys

import numpy as np
import pytest

import gudni_amd as G
from gudni_amd import scenes as S
from oracle import oracle as O
from tests import helpers as H
from tests.test_gpu_binning import compare_binning
from tests.test_gpu_parity import dump_parity

pytestmark = pytest.mark.gpu

N_CASES = 16


def say(msg):
    print(f"[gpu-random-records] {msg}", flush=True)
    sys.stdout.flush()


def case(i):
    rng = np.random.default_rng(7000 + i)
    size = int(rng.choice([64, 128, 200, 256, 384]))
    if i % 2 == 0:
        scene = S.batch_scene(size, int(rng.integers(5, 250)), 400 + i, gradient_share=int(rng.integers(0, 4)),
                              picture_share=int(rng.integers(0, 3)))
    else:
        scene = S.blobs_scaled(size, int(rng.integers(20, 400)), 500 + i, glyph_share=float(rng.uniform(0.05, 0.4)))
    spec = G.default_spec(max_tile_size=int(rng.choice([16, 32, 64, 128, 256])),
                          sort_radix_bits=int(rng.integers(0, 6)))
    return scene, spec


@pytest.mark.parametrize("i", range(N_CASES))
def test_random_records_match_oracle(i):
    scene, spec = case(i)
    say(f"case {i}: {scene.name} {scene.width}x{scene.height} {scene.n_shapes} shapes, tile "
        f"{spec.max_tile_size} radix {spec.sort_radix_bits}")
    _, geo, jobs = G.prepare(scene, spec)
    oj = O.JobList(H.oracle_spec(spec), O.Geometry(scene), scene.width, scene.height)
    assert oj.n_jobs == jobs.n_jobs
    np.testing.assert_array_equal(oj.tiles, jobs.tiles, err_msg="tile records")
    np.testing.assert_array_equal(oj.shape_refs, jobs.shape_refs, err_msg="shape references")
    r = G.Rasterizer(spec)
    try:
        compare_binning(r, scene, spec)
        dump_parity(r, scene, lambda n: sorted({0, n - 1}), cap=256, spec=spec)
    finally:
        r.close()
