"""smcrt_run_device_noalloc: the packet loop with no device allocation, every buffer beyond the
caller's grid in the library's file-scope scratch array.  Against smcrt_run_device on the same
packets (identical counters, fluence to FP64 summation order), on two calls in a row with the
second one smaller (fewer packets, a smaller scene and grid), and without changing the device's
free memory."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from signedmcrt_amd import _abi as A
from signedmcrt_amd import scenes
from signedmcrt_amd.engine import Engine, run_opts

pytestmark = pytest.mark.gpu

KEYS = ("launched", "absorbed", "escaped_r", "escaped_t", "killed_surface", "killed_bounce", "scatters",
        "ev_sweeps", "ev_steps", "ev_crossings", "ev_interfaces", "ev_fresnel", "err_unreachable", "err_nan")


def _noalloc(eng, n, d, seed, first=0, stats=True, **shape):
    o = run_opts(n, seed, first, **shape)
    st = A.Stats()
    A.check(A.lib().smcrt_run_device_noalloc(eng.handle, C.byref(o), C.c_void_p(d.data_ptr()),
                                             C.byref(st) if stats else None, None))
    return st.as_dict()


def _compare(eng, n, seed, first=0):
    import torch
    g = eng.scene.grid
    nv = g.nxg * g.nyg * g.nzg
    a = torch.zeros(nv, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a)
    sa = _noalloc(eng, n, a, seed, first)
    sb = eng.run_device(n, b.data_ptr(), seed=seed, first_packet=first)
    torch.cuda.synchronize()
    for k in KEYS:
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    x, y = a.cpu().numpy(), b.cpu().numpy()
    assert abs(x.sum() / y.sum() - 1) < 1e-10
    m = y >= 1e-3 * y.max()
    assert np.max(np.abs(x[m] / y[m] - 1)) < 1e-9
    assert abs(sa["tally_sum"] / y.sum() - 1) < 1e-10
    return sa


def test_noalloc_matches_run_device_twice(gpu):
    import torch
    big, _ = scenes.build("skin_vessels", nxyz=64)
    small, _ = scenes.build("slab", nxyz=40)
    e1, e2 = Engine(big), Engine(small)
    _compare(e1, 400_000, seed=5)                # first call: larger scene, grid and launch
    _compare(e2, 50_000, seed=6, first=123)      # second call: smaller everything
    _compare(e1, 200_000, seed=7, first=999)
    # no call allocates: the device's free memory is unchanged by further allocation-free calls
    d = torch.zeros(small.grid.nxg * small.grid.nyg * small.grid.nzg, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    for k in range(3):
        _noalloc(e2, 20_000, d, seed=k)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free1


def test_noalloc_without_stats_is_asynchronous_and_equal(gpu):
    """stats == NULL: the call only enqueues; after a synchronize the grid equals the waited call's."""
    import torch
    sc, _ = scenes.build("refracting_sphere", nxyz=50)
    e = Engine(sc)
    nv = 50 ** 3
    a = torch.zeros(nv, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a)
    _noalloc(e, 100_000, a, seed=3, stats=False)
    _noalloc(e, 100_000, b, seed=3, stats=True)
    torch.cuda.synchronize()
    x, y = a.cpu().numpy(), b.cpu().numpy()
    assert abs(x.sum() / y.sum() - 1) < 1e-10


def test_noalloc_limits(gpu):
    """Tracing is rejected (it needs a buffer); a grid whose partial copies do not fit in the scratch
    is tallied directly and still matches."""
    import torch
    sc, _ = scenes.build("slab", nxyz=40)
    e = Engine(sc)
    d = torch.zeros(40 ** 3, dtype=torch.float64, device="cuda")
    o = run_opts(1000, 1, trace_packets=4, trace_events=4)
    rc = A.lib().smcrt_run_device_noalloc(e.handle, C.byref(o), C.c_void_p(d.data_ptr()), None, None)
    assert rc == A.EINVAL and "trac" in A.last_error()
    big, _ = scenes.build("skin_vessels", nxyz=256)  # 128 MiB per partial copy > the scratch
    assert 256 ** 3 * 8 * 2 > A.SCRATCH_BYTES
    _compare(Engine(big), 300_000, seed=11)
