"""The seeded scene generators: records the C ABI accepts, and reproducible
output for a seed."""
import numpy as np
import pytest

from gudni_amd import scenes as S


def _check_gradients(sc):
    g = np.asarray(sc.gradients, np.float32).reshape(-1, 28)
    iv = g.view(np.int32)
    assert np.isin(iv[:, 0], (0, 1)).all(), "kind must be linear (0) or radial (1)"
    assert ((iv[:, 1] >= 1) & (iv[:, 1] <= 4)).all(), "1..4 stops"
    for row, n in zip(g, iv[:, 1]):
        offs = row[8:8 + n]
        assert (np.diff(offs) >= 0).all() and offs[0] == 0.0 and offs[-1] == 1.0
        assert ((row[12:12 + 4 * n] >= 0) & (row[12:12 + 4 * n] <= 1)).all()
    # every gradient substance names a record of the table
    tags = np.asarray(sc.tags, np.uint64)
    subs = np.asarray(sc.substances, np.float32).view(np.uint32)
    sid = (tags & np.uint64(S.TAG_ID_MASK)).astype(np.int64)
    kinds = (tags >> np.uint64(62)).astype(np.int64)
    for s, k in zip(sid, kinds):
        if k == 3:
            assert subs[s, 0] < len(g)


@pytest.mark.parametrize("make", [
    lambda: S.batch_scene(256, 300, 5),
    lambda: S.batch_scene(512, 800, 11, gradient_share=5),
    lambda: S.blobs_4096(n_blobs=3000, n_glyphs=300, size=1024),
])
def test_gradient_tables_are_valid(make):
    sc = make()
    assert len(sc.gradients) > 0
    _check_gradients(sc)


def test_gradient_record_rejects_bad_stops():
    with pytest.raises(ValueError):
        S.gradient_record(0, (0, 0), (1, 1), [0, .25, .5, .75, 1.0], np.zeros((5, 4)))
    with pytest.raises(ValueError):
        S.gradient_record(2, (0, 0), (1, 1), [0, 1.0], np.zeros((2, 4)))


def test_generators_are_reproducible():
    a, b = S.adjoint_1024(), S.adjoint_1024()
    assert np.array_equal(a.pairs, b.pairs) and np.array_equal(a.substances, b.substances)
    p, q = S.perturb(a, 3), S.perturb(a, 3)
    assert np.array_equal(p.pairs, q.pairs)
    assert not np.array_equal(p.pairs, a.pairs)
    assert a.n_shapes == 2000
