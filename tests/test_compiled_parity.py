"""Oracle parity of the packet kernel compiled at run time for the scene (SMCRT_JIT=1, built for 4
resident waves per SIMD: the kernel the benchmark times) on the slab (1e6-packet config, 100^3
grid), the refracting sphere (200^3) and skin + vessels (256^3) workloads at their own grids, on
a four-polygon version of the polygon workload, and of its adjoint on skin + vessels.

Bars:
  - absorbed / reflected / transmitted / killed fractions: two-sample z within 3;
  - per-voxel fluence, no coarsening: z of the difference of the two batch means (B = 20 batches a
    side, standard errors from the batches) over voxels whose fluence is at least `floor` of the
    maximum; at most 1 % of them beyond |z| = 3 and |mean z| <= 0.1, with >= 1e4 voxels compared.
    A pencil beam's fluence falls by four decades within a few voxels of the axis (466 voxels of
    the slab and 150 of skin + vessels hold >= 1 % of the maximum), so the floor is set per
    workload to bring >= 1e4 voxels in: 1e-3 of the maximum for the slab, 3e-4 for skin +
    vessels, 1e-2 for the refracting sphere's broad beam and the polygons' uniform source.  The
    same test run oracle against oracle (two seeds, these batch sizes) gives 0.2-0.5 % beyond
    |z| = 3 and |mean z| <= 0.035;
  - the per-packet fluence standard deviation (median over the same voxels of sd_gpu/sd_oracle,
    each from its batches) within 3 % of 1: the two estimators have the same variance, so
    packets/s compares the two at equal fluence error (oracle against oracle: within 0.3 %);
  - adjoint: d(R_d)/d(mu_a, mu_s) and d(fluence sum)/d(mu_a, mu_s) of the five tissue layers
    against central differences of the compiled forward kernel over the same packets (common
    random numbers) and against the oracle's score-function gradients, |z| < 4 on batch means.
"""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as O
from signedmcrt_amd import scenes
from stats_util import two_sample_z

pytestmark = pytest.mark.gpu

B = 20
CASES = {
    # workload: (GPU packets per batch, oracle packets per batch, voxel floor)
    "slab": (500_000, 50_000, 1e-3),
    "refracting_sphere": (250_000, 25_000, 1e-2),
    "skin_vessels": (1_000_000, 50_000, 3e-4),
    # four polygons of 40 edges (the gathered polygon evaluation, disjoint rectangles) on an
    # 80 x 80 x 10 grid: the oracle evaluates each polygon by brute force
    "polygon_regions": (500_000, 50_000, 1e-2),
}
BUILD_KW = {"polygon_regions": {"n_side": 2, "n_edges": 40, "nxy": 80, "nz": 10}}
SHAPE = {"waves": 4}
FRACTIONS = ("absorbed", "escaped_r", "escaped_t", "killed_surface")


class Moments:
    """Per-voxel running sum and sum of squares of batch tallies / batch packets."""

    def __init__(self):
        self.s = self.s2 = None
        self.k = 0

    def add(self, x):
        self.s = x.copy() if self.s is None else self.s + x
        self.s2 = x * x if self.s2 is None else self.s2 + x * x
        self.k += 1

    def mean_se(self):
        m = self.s / self.k
        var = np.maximum(self.s2 / self.k - m * m, 0.0) * self.k / (self.k - 1)
        return m, np.sqrt(var / self.k)


def _compiled_engine(sc, monkeypatch, tmp_path):
    from signedmcrt_amd.engine import Engine
    monkeypatch.setenv("SMCRT_JIT", "1")
    monkeypatch.setenv("SMCRT_CACHE_DIR", str(tmp_path))
    return Engine(sc)


def _gpu_batches(eng, sc, n, seed):
    import torch
    g = sc.grid
    nv = g.nxg * g.nyg * g.nzg
    d = torch.zeros(nv, dtype=torch.float64, device="cuda")
    s = torch.zeros_like(d)
    s2 = torch.zeros_like(d)
    tot = {k: 0 for k in FRACTIONS + ("launched", "killed_bounce", "err_unreachable", "err_nan")}
    for b in range(B):
        d.zero_()
        st = eng.run_device(n, d.data_ptr(), seed=seed, first_packet=b * n, **SHAPE)
        for k in tot:
            tot[k] += st[k]
        x = d / n
        s += x
        s2 += x * x
    torch.cuda.synchronize()
    mom = Moments()
    mom.s = s.cpu().numpy().reshape(g.shape, order="F")
    mom.s2 = s2.cpu().numpy().reshape(g.shape, order="F")
    mom.k = B
    return mom, tot


def _oracle_batches(sc, n, seed):
    mom = Moments()
    tot = {k: 0 for k in FRACTIONS + ("launched", "killed_bounce", "err_unreachable", "err_nan")}
    for b in range(B):
        r = O.run(sc, n, seed=seed, first_packet=10**9 + b * n)
        for k in tot:
            tot[k] += r["stats"][k]
        mom.add(r["jmean"] / n)
    return mom, tot


@pytest.mark.parametrize("name", list(CASES))
def test_compiled_kernel_matches_oracle(gpu, name, monkeypatch, tmp_path):
    n_gpu, n_orc, floor = CASES[name]
    sc, _ = scenes.build(name, **BUILD_KW.get(name, {}))
    eng = _compiled_engine(sc, monkeypatch, tmp_path)
    mg, tg = _gpu_batches(eng, sc, n_gpu, seed=2025)
    kname = eng.kernel_variant()[0]
    assert "compiled" in kname, kname
    mo, to = _oracle_batches(sc, n_orc, seed=3031)
    for t, n in ((tg, n_gpu), (to, n_orc)):
        assert t["launched"] == B * n
        assert sum(t[k] for k in FRACTIONS) + t["killed_bounce"] == B * n
        assert t["err_unreachable"] == 0 and t["err_nan"] == 0
    for k in FRACTIONS:
        z = two_sample_z(tg[k], tg["launched"], to[k], to["launched"])
        assert abs(z) < 3.0, (name, k, z, tg[k] / tg["launched"], to[k] / to["launched"])
    m1, se1 = mg.mean_se()
    m2, se2 = mo.mean_se()
    pooled = (m1 + m2) / 2
    mask = pooled >= floor * pooled.max()
    den = np.sqrt(se1[mask] ** 2 + se2[mask] ** 2)
    ok = den > 0
    z = (m1[mask][ok] - m2[mask][ok]) / den[ok]
    assert len(z) >= 10_000, len(z)
    frac3 = float(np.mean(np.abs(z) > 3))
    assert frac3 <= 0.01, (name, frac3, np.sort(np.abs(z))[-10:])
    assert abs(float(np.mean(z))) <= 0.1, (name, float(np.mean(z)))
    # per-packet standard deviations: sd of a batch mean x sqrt(batch packets)
    both = (se1[mask] > 0) & (se2[mask] > 0)
    ratio = float(np.median((se1[mask][both] * np.sqrt(n_gpu)) / (se2[mask][both] * np.sqrt(n_orc))))
    assert abs(ratio - 1) <= 0.03, (name, ratio)
    print(f"{name}: {kname} voxels {len(z)} |z|>3 {frac3:.4f} mean z {np.mean(z):+.3f} "
          f"sd ratio {ratio:.4f}", flush=True)


def _grad_batches(run, nb):
    keys = ("d_r_dmua", "d_r_dmus", "d_w_dmua", "d_w_dmus")
    out = {k: [] for k in keys}
    for b in range(nb):
        g = run(b)
        for k in keys:
            out[k].append(np.asarray(g[k] if isinstance(g, dict) else getattr(g, k)))
    return {k: np.array(v) for k, v in out.items()}


def _z(a, b):
    se = np.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    return (a.mean() - b.mean()) / max(se, 1e-300)


def test_compiled_adjoint_skin_vessels(gpu, monkeypatch, tmp_path):
    """The compiled adjoint on skin + vessels (256^3): d(R_d, fluence sum)/d(mu_a, mu_s) of the five
    layers (materials 1-5; 0 is blood, 6 air) against CRN central differences of the compiled
    forward kernel and against the oracle."""
    import torch
    sc, _ = scenes.build("skin_vessels")
    eng = _compiled_engine(sc, monkeypatch, tmp_path)
    layers = [1, 2, 3, 4, 5]
    nb, n = 8, 2_000_000
    g = _grad_batches(lambda b: eng.grad(n, seed=77, first_packet=b * n, **SHAPE), nb)
    assert "compiled" in eng.kernel_variant()[0]
    # CRN central differences, relative step 5 %
    mus0 = np.array([m.mus for m in sc.materials])
    mua0 = np.array([m.mua for m in sc.materials])
    nv = sc.grid.nxg * sc.grid.nyg * sc.grid.nzg
    d = torch.zeros(nv, dtype=torch.float64, device="cuda")
    for which, base in (("mua", mua0), ("mus", mus0)):
        for mat in layers:
            h = 0.05 * base[mat]
            fd = []
            for b in range(nb):
                v = []
                for s in (+1, -1):
                    mus, mua = mus0.copy(), mua0.copy()
                    (mua if which == "mua" else mus)[mat] += s * h
                    eng.set_materials(mus=mus, mua=mua)
                    d.zero_()
                    st = eng.run_device(n, d.data_ptr(), seed=77, first_packet=b * n, **SHAPE)
                    v.append((st["escaped_r"] / n, st["tally_sum"] / n))
                fd.append(((v[0][0] - v[1][0]) / (2 * h), (v[0][1] - v[1][1]) / (2 * h)))
            eng.set_materials(mus=mus0, mua=mua0)
            fd = np.array(fd)
            zr = _z(g[f"d_r_d{which}"][:, mat], fd[:, 0])
            zw = _z(g[f"d_w_d{which}"][:, mat], fd[:, 1])
            assert abs(zr) < 4 and abs(zw) < 4, (which, mat, zr, zw, g[f"d_w_d{which}"][:, mat].mean(), fd[:, 1].mean())
    # the oracle's FP64 score-function gradients (independent packets)
    no = 100_000
    go = _grad_batches(lambda b: O.run(sc, no, seed=78, first_packet=10**9 + b * no, grad=True, tally=False), 10)
    for k in go:
        for mat in layers:
            z = _z(g[k][:, mat], go[k][:, mat])
            assert abs(z) < 4, (k, mat, z, g[k][:, mat].mean(), go[k][:, mat].mean())


def test_lds_tile_at_bench_load(gpu, monkeypatch, tmp_path):
    """The LDS tile sums in FP64.  At and above the benchmark's per-block load (one block per CU:
    ~4e4 packets per block here, 1e5 in the 1e8-packet benchmark launch) the compiled kernel's
    grid equals the direct FP64 tally of the same packets (FLAG_NO_LDS_TILE) to 1e-10 on the total
    and 1e-9 on every voxel holding >= 1e-3 of the maximum, with identical counters.  (An FP32
    tile drifted here by +1.4e-5 on the total: LDS float adds round like register adds, and a
    sequential FP32 sum of 1e4-1e5 chords carries ~1e-5 relative error.)"""
    from signedmcrt_amd import _abi as A
    sc, _ = scenes.build("skin_vessels")
    eng = _compiled_engine(sc, monkeypatch, tmp_path)
    kw = dict(SHAPE, blocks_per_cu=1)
    a = eng.run(10_000_000, seed=19, **kw)
    assert "compiled" in eng.kernel_variant()[0]
    b = eng.run(10_000_000, seed=19, flags=A.FLAG_NO_LDS_TILE, **kw)
    for k in ("absorbed", "escaped_r", "escaped_t", "killed_surface", "scatters", "ev_crossings"):
        assert a.stats[k] == b.stats[k], k
    assert a.stats["ev_atomics"] < b.stats["ev_atomics"]  # the tile took part
    assert abs(a.jmean.sum() / b.jmean.sum() - 1) < 1e-10
    m = b.jmean >= 1e-3 * b.jmean.max()
    assert np.max(np.abs(a.jmean[m] / b.jmean[m] - 1)) < 1e-9


@pytest.mark.parametrize("jit", ["0", "1"])
def test_clear_slab_fresnel_on_gpu(gpu, jit, monkeypatch, tmp_path):
    """The interface table and Fresnel events on the GPU (the library's interpreting kernel and
    the compiled one) against closed forms: a clear n = 1.5 slab in air under a collimated beam
    (killed = R + (1 - R) R, transmitted (1 - R)^2, R the textbook Fresnel reflectance), the
    60-degree grazing-crossing kill, total internal reflection from inside at 45 degrees and
    R(30 degrees, 1.5 -> 1) from inside."""
    import math
    from pin_scenes import clear_slab, clear_slab_expected, fresnel_textbook
    from signedmcrt_amd.engine import Engine
    monkeypatch.setenv("SMCRT_JIT", jit)
    monkeypatch.setenv("SMCRT_CACHE_DIR", str(tmp_path))
    n = 1_000_000
    cases = [(clear_slab(30.0), clear_slab_expected(30.0)), (clear_slab(45.0), clear_slab_expected(45.0)),
             (clear_slab(60.0), {"killed_surface": 1.0}), (clear_slab(45.0, inside=True), {"killed_surface": 1.0})]
    R = fresnel_textbook(math.radians(30.0), 1.5, 1.0)
    cases.append((clear_slab(30.0, inside=True), {"killed_surface": R, "escaped_t": 1 - R}))
    for sc, want in cases:
        eng = Engine(sc)
        st = eng.run(n, seed=31, tally=False, **SHAPE).stats
        assert ("compiled" in eng.kernel_variant()[0]) == (jit == "1")
        assert st["launched"] == n and st["err_unreachable"] == 0 and st["err_nan"] == 0
        for k, w in want.items():
            sd = math.sqrt(max(w * (1 - w), 1.0 / n) / n)
            assert abs(st[k] / n - w) <= 4 * sd, (k, st[k] / n, w)
