"""Launch-shape autotuner: the search (CPU, on a synthetic cost), the cache, and on the GPU that
every shape transports the same packets and the tuned shape is no slower than the default."""
from __future__ import annotations

import itertools
import json

import pytest

from signedmcrt_amd import autotune, scenes
from signedmcrt_amd import _abi as A


def _separable_cost(sh):
    # a cost with one minimum per knob plus a coupling term that coordinate descent must handle
    c = {64: 3.0, 128: 1.0, 256: 2.0}[sh["block"]]
    c += abs(sh["chunk"] - 256) / 128.0
    c += abs(sh["walk_cap"] - 8) / 4.0 + (0.5 if sh["walk_cap"] == 8 and sh["chunk"] == 64 else 0.0)
    c += {1: 0.3, 8: 0.0, 16: 0.2, 32: 0.4}[sh["refill_min"]]
    c += 0.0 if sh["tile"] else 1.0
    c += {2: 0.1, 3: 0.0, 4: 0.05}[sh["waves"]]
    return c


def test_coordinate_descent_finds_the_grid_minimum():
    best, table = autotune.coordinate_descent(_separable_cost, autotune.SPACE, autotune.DEFAULT)
    brute = min((dict(zip(autotune.SPACE, v)) for v in itertools.product(*autotune.SPACE.values())),
                key=_separable_cost)
    assert _separable_cost(best) == pytest.approx(_separable_cost(brute))
    assert best == brute
    # far fewer probes than the full product (3*4*5*4*2*3 = 1440)
    assert len(table) < 40
    assert table[0][0] == autotune.DEFAULT


def test_small_gains_do_not_move_the_shape():
    """A point 1 % cheaper than the default (within timing noise) is not taken with
    min_gain = 2 %; one 5 % cheaper is."""
    def cost(sh):
        c = 100.0
        if sh["block"] == 128:
            c -= 1.0
        if sh["walk_cap"] == 8:
            c -= 5.0
        return c
    best, _ = autotune.coordinate_descent(cost, autotune.SPACE, autotune.DEFAULT, min_gain=0.02)
    assert best["block"] == autotune.DEFAULT["block"] and best["walk_cap"] == 8
    best0, _ = autotune.coordinate_descent(cost, autotune.SPACE, autotune.DEFAULT, min_gain=0.0)
    assert best0["block"] == 128


def test_run_kwargs_maps_tile_to_flag():
    kw = autotune.run_kwargs({"block": 128, "chunk": 256, "walk_cap": 2, "refill_min": 8, "tile": False})
    assert kw == {"block": 128, "chunk": 256, "walk_cap": 2, "refill_min": 8, "flags": A.FLAG_NO_LDS_TILE}
    assert autotune.run_kwargs(autotune.DEFAULT)["flags"] == 0


def test_run_kwargs_maps_grid_copies_to_flag():
    assert autotune.run_kwargs({"copies": 1})["flags"] == 0
    assert autotune.run_kwargs({"copies": 8})["flags"] == 4 << A.FLAG_GRID_COPIES_SHIFT
    assert autotune.run_kwargs({"copies": 2, "tally": "direct"})["flags"] == (2 << A.FLAG_GRID_COPIES_SHIFT) | A.FLAG_DIRECT_TALLY
    with pytest.raises(ValueError):
        A.flag_grid_copies(3)


def test_explicit_grid_copies_replace_the_tuned_ones():
    from signedmcrt_amd.engine import Engine
    e = Engine.__new__(Engine)
    e.shape = {"copies": 8, "tally": "direct"}
    flags, kw = e._shape(A.flag_grid_copies(2), {})
    assert flags & A.FLAG_GRID_COPIES_MASK == A.flag_grid_copies(2)
    assert flags & A.FLAG_DIRECT_TALLY
    flags, kw = e._shape(0, {})
    assert flags & A.FLAG_GRID_COPIES_MASK == A.flag_grid_copies(8)


def test_tuned_shape_runs_a_batch():
    """bench.py's batched workload hands part of the tuner's shape to Batch, which builds its
    launch options with engine.run_opts: the knobs batch_shape keeps (block size, walk cap) map to
    accepted keywords, together with Batch.fastest's candidates (waves, refill policy); chunk and
    refill policy stay unset (0) for the library's batch defaults."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    from signedmcrt_amd.engine import run_opts
    shape = {k: v[-1] for k, v in autotune.SPACE.items()}
    shape["adj_waves"] = 4
    kept = bench.batch_shape(shape)
    assert set(kept) == {"block", "walk_cap"}
    o = run_opts(0, **kept)
    assert o.block_size == shape["block"] and o.walk_cap == shape["walk_cap"]
    assert o.chunk == 0 and o.refill_min == 0
    o = run_opts(0, **dict(kept, waves=4, refill_min=16))
    assert o.waves == 4 and o.refill_min == 16


def test_fingerprint_tracks_the_scene():
    a, _ = scenes.build("slab")
    b, _ = scenes.build("slab")
    c, _ = scenes.build("slab", mua=0.2)
    d, _ = scenes.build("slab", nxyz=64)
    assert autotune.fingerprint(a) == autotune.fingerprint(b)
    assert autotune.fingerprint(a) != autotune.fingerprint(c)
    assert autotune.fingerprint(a) != autotune.fingerprint(d)


@pytest.mark.gpu
def test_shapes_transport_identical_packets(gpu):
    """Regeneration policy, chunk, walk cap and block size change only the schedule."""
    from signedmcrt_amd.engine import Engine
    sc, _ = scenes.build("refracting_sphere", nxyz=64)
    eng = Engine(sc)
    ref = eng.run(60_000, seed=5).stats
    for sh in ({"refill_min": 32}, {"refill_min": 64, "chunk": 512}, {"walk_cap": 1, "block": 64},
               {"walk_cap": 16, "chunk": 128, "refill_min": 8}, {"waves": 2}, {"waves": 3, "block": 128}, {"waves": 4}):
        st = eng.run(60_000, seed=5, **sh).stats
        for k in autotune.INVARIANT:
            assert st[k] == ref[k], (sh, k)


@pytest.mark.gpu
def test_tune_on_gpu_and_cache(gpu, tmp_path):
    from signedmcrt_amd.engine import Engine
    sc, _ = scenes.build("slab", nxyz=64)
    eng = Engine(sc)
    cache = str(tmp_path / "tune.json")
    res = eng.tune(n_probe=300_000, reps=2, sweeps=1, cache_path=cache)
    assert not res.cached
    assert set(res.shape) == set(autotune.SPACE) | {"adj_waves"}
    assert res.shape["adj_waves"] in autotune.SPACE["waves"]
    # the default point is measured too; the chosen point is the best measured one
    assert res.kernel_ms <= res.default_ms
    # a knob moves only for a gain above min_gain (2 %), so the chosen point is within 2 % of
    # the best measured one
    assert res.kernel_ms <= min(c for _, c in res.table) / 0.98
    data = json.load(open(cache))
    assert len(data) == 1 and list(data.values())[0]["shape"] == res.shape
    again = autotune.tune(eng, cache_path=cache)
    assert again.cached and again.shape == res.shape
    # the engine now runs with the tuned shape and still transports the same packets
    r1 = eng.run(40_000, seed=9).stats
    r2 = Engine(sc).run(40_000, seed=9).stats
    for k in autotune.INVARIANT:
        assert r1[k] == r2[k], k
