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
        # the standalone run sums its hot tile in LDS first: equal to FP64 summation order
        np.testing.assert_allclose(jm, r.jmean, rtol=1e-9, atol=1e-13)


def test_compiled_batch_reproduces_compiled_runs(gpu, monkeypatch, tmp_path):
    """SMCRT_JIT=1: the geometry groups run the scenes' run-time compiled kernels for several sets
    (set lookup and per-lane outcome counters), the standalone runs the single-set compiled
    kernels: same packets, same transport."""
    monkeypatch.setenv("SMCRT_JIT", "1")
    monkeypatch.setenv("SMCRT_CACHE_DIR", str(tmp_path))
    engines = _engines()
    rng = np.random.default_rng(11)
    sets = _sets(engines, rng, 8)
    b = Batch(engines, sets, seed=77)
    b.run()
    for eng in engines:
        name = eng.kernel_variant()[0]
        assert "compiled" in name and "failed" not in name, name
    for i, s in enumerate(sets):
        jm, st = b.read(i)
        eng = engines[s.scene]
        eng.set_materials(mus=np.array(s.mus), mua=np.array(s.mua))
        r = eng.run(s.n_packets, seed=77, first_packet=s.first_packet)
        for k in OUT:
            assert st[k] == r.stats[k], (i, k, st[k], r.stats[k])
        np.testing.assert_allclose(jm, r.jmean, rtol=1e-9, atol=1e-13)


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


def test_fastest_wave_count_keeps_results(gpu, monkeypatch, tmp_path):
    """Batch.fastest keeps the batch at the faster of 3 and 4 waves per SIMD (the compiled kernels
    of several sets): its counters equal those of the batch at the other wave count, its grids
    within FP64 summation order."""
    monkeypatch.setenv("SMCRT_JIT", "1")
    monkeypatch.setenv("SMCRT_CACHE_DIR", str(tmp_path))
    engines = _engines()
    sets = _sets(engines, np.random.default_rng(9), 6)
    b = Batch.fastest(engines, sets, seed=31)
    assert b.shape["waves"] in (3, 4)
    assert b.replay_ms() > 0.0
    other = Batch(engines, sets, seed=31, waves=7 - b.shape["waves"])
    other.run()
    for i in range(len(sets)):
        jm, st = b.read(i)
        jo, so = other.read(i)
        for k in OUT:
            assert st[k] == so[k], (i, k, st[k], so[k])
        np.testing.assert_allclose(jm, jo, rtol=1e-9, atol=1e-12)


# ---- the batched workload (256 seeded property sets) against the oracle -----------------------
def _workload_sets(n_sets, cap=None):
    """Property sets of the batched benchmark (bench.batched_sets): even sets on the slab, odd on
    the refracting sphere (64^3 grids), the drawn mu_a, mu_s, g and n on the scattering medium
    (the slab; the sphere's cube), budgets log-uniform 1e5-1e7 (capped at `cap`)."""
    import bench
    engines = _engines()
    medium = [0, 1]
    sets, first = [], 0
    for geo, mua, mus, g, n, budget in bench.batched_sets(n_sets):
        mats = engines[geo].scene.materials
        MUS, MUA = [m.mus for m in mats], [m.mua for m in mats]
        HG, NN = [m.hgg for m in mats], [m.n for m in mats]
        k = medium[geo]
        MUS[k], MUA[k], HG[k], NN[k] = mus, mua, g, n
        budget = min(budget, cap) if cap else budget
        sets.append(BatchSet(geo, MUS, MUA, budget, first_packet=first, hgg=HG, n=NN))
        first += budget
    return engines, sets


def _oracle_scene(engines, s):
    """The set's scene for the oracle: a copy of its geometry with the set's materials."""
    import copy
    from signedmcrt_amd.scene import Scene
    sc = engines[s.scene].scene
    tmp = Scene(copy.deepcopy(sc.array), sc.grid, sc.source)
    for i, m in enumerate(tmp.materials):
        m.mus, m.mua, m.hgg, m.n = s.mus[i], s.mua[i], s.hgg[i], s.n[i]
    return Scene(tmp.array, tmp.grid, tmp.source)


def test_batched_sets_match_oracle(gpu):
    """The first 16 sets of the batched workload (8 per geometry, per-set mu_a, mu_s, g, n), run in
    one batch (budgets capped at 1e6), against the oracle on independent packets of each set:
    absorbed / reflected / transmitted / killed fractions and the fluence total per packet.  Over
    the 80 comparisons: every |z| < 4, at most 2 beyond 3 (0.2 expected), and the sum of z^2
    consistent with chi^2(80) (p > 1e-3)."""
    import pyoracle as O
    from scipy import stats as S
    from stats_util import two_sample_z
    engines, sets = _workload_sets(16, cap=1_000_000)
    b = Batch(engines, sets, seed=424242)
    b.run()
    zs = []
    for i, s in enumerate(sets):
        jm, st = b.read(i)
        ng = st["launched"]
        assert ng == s.n_packets
        assert st["absorbed"] + st["escaped_r"] + st["escaped_t"] + st["killed_surface"] + st["killed_bounce"] == ng
        assert st["err_unreachable"] == 0 and st["err_nan"] == 0
        sc = _oracle_scene(engines, s)
        nb, B = 1500, 20  # 20 batches: the per-packet sd enters the total's z with 19 dof
        ko = {k: 0 for k in ("absorbed", "escaped_r", "escaped_t", "killed_surface")}
        tot = []
        for j in range(B):
            r = O.run(sc, nb, seed=777, first_packet=10**12 + i * 10**7 + j * nb)
            for k in ko:
                ko[k] += r["stats"][k]
            tot.append(r["jmean"].sum() / nb)
        for k in ko:
            zs.append(two_sample_z(st[k], ng, ko[k], B * nb))
        # fluence total per packet: the oracle's per-packet sd for both sides (same estimator)
        sd = np.std(tot, ddof=1) * np.sqrt(nb)
        zs.append((jm.sum() / ng - np.mean(tot)) / (sd * np.sqrt(1.0 / ng + 1.0 / (B * nb))))
    zs = np.array(zs)
    assert np.all(np.abs(zs) < 4), zs
    assert np.sum(np.abs(zs) > 3) <= 2, zs
    assert S.chi2.sf(np.sum(zs ** 2), len(zs)) > 1e-3, (np.sum(zs ** 2), len(zs))


def test_batched_workload_full_size(gpu):
    """The whole batched workload: 256 sets, full budgets (5.9e8 packets, one graph replay).  Per
    set: every budgeted packet launched and ended in exactly one class; per geometry group: no
    error counters; and the fluence of each set is finite and non-negative."""
    engines, sets = _workload_sets(256)
    b = Batch(engines, sets, seed=123456789)
    b.run()
    total = 0
    for i, s in enumerate(sets):
        jm, st = b.read(i)
        assert st["launched"] == s.n_packets, (i, st["launched"], s.n_packets)
        assert st["absorbed"] + st["escaped_r"] + st["escaped_t"] + st["killed_surface"] + st["killed_bounce"] == s.n_packets
        assert st["err_unreachable"] == 0 and st["err_nan"] == 0
        assert np.all(np.isfinite(jm)) and jm.min() >= 0.0 and jm.sum() > 0.0
        total += st["launched"]
    assert total == sum(s.n_packets for s in sets) > 5e8
    assert st["kernel_ms"] > 0
