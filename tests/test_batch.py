"""Batched launch under a HIP graph: each set reproduces a standalone run of the same packets
(same random streams), and graph replays repeat the same results."""
from __future__ import annotations

import numpy as np
import pytest

from signedmcrt_amd import scenes
from signedmcrt_amd.batch import Batch, BatchSet
from signedmcrt_amd.engine import Engine

pytestmark = pytest.mark.gpu

OUT = ("launched", "absorbed", "escaped_r", "escaped_t", "killed_surface", "killed_bounce", "scatters")


def _engines():
    slab, _ = scenes.build("slab", nxyz=64)
    sph, _ = scenes.build("refracting_sphere", nxyz=64)
    return [Engine(slab), Engine(sph)]


def _sets(engines, rng, k):
    sets = []
    for i in range(k):
        sc = i % 2
        base = engines[sc].scene.materials
        mus = [m.mus * rng.uniform(0.5, 1.5) if m.mus > 0 else 0.0 for m in base]
        mua = [m.mua * rng.uniform(0.5, 2.0) if m.mua > 0 else 0.0 for m in base]
        n = int(rng.integers(2_000, 60_000))
        sets.append(BatchSet(sc, mus, mua, n, first_packet=int(rng.integers(0, 1 << 40))))
    return sets


def test_batch_sets_reproduce_standalone_runs(gpu):
    engines = _engines()
    rng = np.random.default_rng(7)
    sets = _sets(engines, rng, 12)
    b = Batch(engines, sets, seed=99)
    b.run()
    for i, s in enumerate(sets):
        jm, st = b.read(i)
        eng = engines[s.scene]
        eng.set_materials(mus=np.array(s.mus), mua=np.array(s.mua))
        r = eng.run(s.n_packets, seed=99, first_packet=s.first_packet)
        for k in OUT:
            assert st[k] == r.stats[k], (i, k, st[k], r.stats[k])
        # the standalone run accumulates its hot tile in FP32 LDS: equal to ~1e-6
        np.testing.assert_allclose(jm, r.jmean, rtol=2e-6, atol=1e-9)


def test_graph_replays_repeat(gpu):
    engines = _engines()
    rng = np.random.default_rng(8)
    sets = _sets(engines, rng, 6)
    b = Batch(engines, sets, seed=5)
    first = []
    b.run()
    for i in range(len(sets)):
        first.append(b.read(i))
    for rep in range(3):
        # eager work between replays on the same device
        engines[0].run(5_000, seed=rep)
        b.run()
        for i in range(len(sets)):
            jm, st = b.read(i)
            for k in OUT:
                assert st[k] == first[i][1][k], (rep, i, k)
            np.testing.assert_allclose(jm, first[i][0], rtol=1e-9, atol=1e-12)
