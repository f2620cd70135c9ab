"""Binned tally (SMCRT_FLAG_BINNED_TALLY: chords outside the LDS tile staged per grid bin in LDS,
written to device pages and summed per bin after the launch) against the direct FP64 atomics
(SMCRT_FLAG_DIRECT_TALLY) on the same packets.  Transport is identical, so every counter must be
equal and the fluence equal up to FP64 summation order: per voxel within 1e-9 relative (a voxel
sums at most ~1e6 FP32 chords, each rounding of an FP64 partial sum is 1.1e-16 relative).

The stage is built into the kernels compiled at run time for the scene (SMCRT_JIT=1 here; the
library's kernels always take the atomics).  Covers a chunk pool too small for the launch (blocks
run out of chunks and their later entries take the atomics), consecutive launches, and the
automatic choice (binned for the refracting sphere, whose LDS tile holds ~4 % of the path length).
The tally-atomics counter differs between the modes and, in the binned one, from run to run (how
many entries find their slot full depends on the wavefronts' timing); every transport counter
must not.
"""
from __future__ import annotations

import numpy as np
import pytest

from signedmcrt_amd import _abi as A
from signedmcrt_amd import scenes

pytestmark = pytest.mark.gpu

COUNTERS = ("launched", "absorbed", "escaped_r", "escaped_t", "killed_surface", "killed_bounce", "scatters",
            "ev_sweeps", "ev_steps", "ev_crossings", "ev_interfaces", "err_unreachable", "err_nan")


@pytest.fixture(scope="module")
def jit_cache(tmp_path_factory):
    return str(tmp_path_factory.mktemp("jit"))


def _engine(name, monkeypatch, cache):
    from signedmcrt_amd.engine import Engine
    monkeypatch.setenv("SMCRT_JIT", "1")
    monkeypatch.setenv("SMCRT_CACHE_DIR", cache)
    sc, _ = scenes.build(name)
    return sc, Engine(sc)


def _run(eng, sc, n, flags):
    import torch
    g = sc.grid
    d = torch.zeros(g.nxg * g.nyg * g.nzg, dtype=torch.float64, device="cuda")
    st = eng.run_device(n, d.data_ptr(), seed=4242, first_packet=7 * n, flags=flags)
    torch.cuda.synchronize()
    return st, d.cpu().numpy()


def _same(sd, gd, sb, gb, name):
    for k in COUNTERS:
        assert sd[k] == sb[k], (name, k, sd[k], sb[k])
    assert gd.sum() > 0
    err = np.abs(gb - gd)
    assert float(err.max()) <= 1e-9 * float(np.abs(gd).max()), (name, float(err.max()))
    assert np.all(err <= 1e-9 * np.abs(gd) + 1e-12 * float(np.abs(gd).max())), name
    assert abs(gb.sum() - gd.sum()) <= 1e-10 * gd.sum()


@pytest.mark.parametrize("name,n", [("slab", 400_000), ("refracting_sphere", 300_000), ("skin_vessels", 500_000)])
def test_binned_matches_direct(gpu, name, n, monkeypatch, jit_cache):
    sc, eng = _engine(name, monkeypatch, jit_cache)
    sd, gd = _run(eng, sc, n, A.FLAG_DIRECT_TALLY)
    sb, gb = _run(eng, sc, n, A.FLAG_BINNED_TALLY)
    assert "compiled" in eng.kernel_variant()[0]
    assert sb["ev_atomics"] < sd["ev_atomics"]  # most entries went through the stage
    _same(sd, gd, sb, gb, name)


def test_binned_with_a_short_chunk_pool(gpu, monkeypatch, jit_cache):
    """64 MiB of chunks for ~1e8 staged entries: blocks run out of chunks early and the rest of
    their entries take the atomics; with no pool at all every entry does.  Same result."""
    sc, eng = _engine("refracting_sphere", monkeypatch, jit_cache)
    sd, gd = _run(eng, sc, 400_000, A.FLAG_DIRECT_TALLY)
    monkeypatch.setenv("SMCRT_TALLY_POOL_MB", "64")
    sb, gb = _run(eng, sc, 400_000, A.FLAG_BINNED_TALLY)
    monkeypatch.setenv("SMCRT_TALLY_POOL_MB", "0")
    s0, g0 = _run(eng, sc, 400_000, A.FLAG_BINNED_TALLY)  # no chunks at all
    _same(sd, gd, sb, gb, "short pool")
    _same(sd, gd, s0, g0, "no pool")


def test_binned_consecutive_launches_and_auto_choice(gpu, monkeypatch, jit_cache):
    """Two launches in a row on one scene (the stage and chunks are reused: a second launch must
    not see the first one's chunks), and the default choice equals both explicit modes."""
    sc, eng = _engine("refracting_sphere", monkeypatch, jit_cache)
    sd, gd = _run(eng, sc, 200_000, A.FLAG_DIRECT_TALLY)
    sb1, gb1 = _run(eng, sc, 200_000, A.FLAG_BINNED_TALLY)
    sb2, gb2 = _run(eng, sc, 200_000, A.FLAG_BINNED_TALLY)
    sa, ga = _run(eng, sc, 200_000, 0)
    # the default chose the stage (how many entries find a slot full, and so take an atomic,
    # depends on the wavefronts' timing: only the transport counters and the sums are exact)
    assert sa["ev_atomics"] < sd["ev_atomics"] / 2
    _same(sd, gd, sb1, gb1, "first")
    _same(sd, gd, sb2, gb2, "second")
    _same(sd, gd, sa, ga, "auto")
