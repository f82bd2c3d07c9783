"""The autotuner's coordinate search (host logic only, a synthetic timing
function): the tile size is chosen at the default radix, then the radix at
that tile size, and the fastest configuration seen wins."""
from gudni_amd import autotune


def test_search_tile_then_radix():
    calls = []
    cost = {(32, 0): 5.0, (64, 0): 6.0, (128, 0): 7.0, (256, 0): 8.0,
            (32, 3): 4.0, (32, 4): 4.5, (32, 5): 9.0}

    def time_fn(t, r):
        calls.append((t, r))
        return cost[(t, r)]

    tile, radix, timings = autotune.tune_frame(time_fn, default_radix=0)
    assert (tile, radix) == (32, 3)
    # every tile at radix 0, then the other radices at tile 32, each once
    assert calls == [(32, 0), (64, 0), (128, 0), (256, 0), (32, 3), (32, 4), (32, 5)]
    assert timings["32,3"] == 4.0


def test_search_keeps_default_radix_when_fastest():
    def time_fn(t, r):
        return (0 if t == 64 else 1) + (0 if r == 4 else 1)

    tile, radix, _ = autotune.search(time_fn, (32, 64, 128), (0, 3, 5), default_radix=4)
    assert (tile, radix) == (64, 4)


def test_fixed_tile_and_radix_lists():
    tile, radix, timings = autotune.tune_frame(lambda t, r: 1.0, tiles=(64,), radices=(5,), default_radix=5)
    assert (tile, radix) == (64, 5) and list(timings) == ["64,5"]
