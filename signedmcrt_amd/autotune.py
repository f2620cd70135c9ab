"""Launch-shape autotuner for the packet kernel (no reference counterpart: the reference runs
OpenMP static chunks, src/kernelsMod.f90:239-269).

Knobs (`SPACE`):
  block       threads per block of the persistent grid (64..256; resident blocks follow from
              the occupancy query)
  chunk       packet ids a wavefront claims per refill; chunk/64 = packets per lane per claim
  walk_cap    voxel chords a lane walks per kernel-loop iteration before it sweeps again
  refill_min  regeneration policy: a wavefront refills its idle lanes once at least this many
              are idle (1 = at once)
  tile        LDS-privatised FP64 tally tile on/off
  tally       where chords outside the tile go: "direct" (FP64 atomics), "binned" (staged per grid
              bin in LDS, written to device pages, summed per bin after the launch) or "auto"
              (the library's choice from the scene's pilot run)
  copies      FP64 partial grids per XCD the tally's global atomics spread over (1, 2, 4, 8):
              more copies, less contention on hot voxels, one grid of memory and a clear and a
              sum per launch each
  waves       resident wavefronts per SIMD the kernel is built for (2, 3 or 4: fewer registers;
              4 only for the run-time compiled kernels, the library's run 3 instead)

Every probe launches the same packets (fixed seed and first packet id).  Each packet draws from
its own (seed, id) random stream, so the transported work is identical for every shape: the
timing difference belongs to the shape alone, and the counters are checked equal across probes.
Search: coordinate descent from the library defaults (two sweeps), best-of-`reps` kernel time per
point.  Results are cached in a JSON file keyed by scene fingerprint, device name and library ABI
version, so a scene is tuned once per device.

    shape = autotune.tune(engine)            # dict of run keywords
    engine.run(n, **shape)                   # or Engine(scene, shape=shape)
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from . import _abi as A

# knobs in the order coordinate descent visits them: resident waves first (the register budget
# changes the kernel's code, and with it which values of the other knobs are best)
SPACE: Dict[str, Tuple] = {
    "waves": (2, 3, 4),
    "block": (64, 128, 256),
    "tile": (True, False),
    "tally": ("auto", "direct", "binned"),
    "copies": (1, 2, 4, 8),
    "chunk": (64, 128, 256, 512),
    "walk_cap": (1, 2, 4, 8, 16),
    "refill_min": (1, 8, 16, 32),
}
DEFAULT = {"block": 256, "chunk": 64, "walk_cap": 4, "refill_min": 1, "tile": True, "tally": "auto", "copies": 1,
           "waves": 3}
# counters that must not change with the launch shape
INVARIANT = ("launched", "absorbed", "escaped_r", "escaped_t", "killed_surface", "killed_bounce",
             "scatters", "ev_steps", "ev_crossings", "ev_interfaces")


def run_kwargs(shape: Dict) -> Dict:
    """Engine.run / run_device keywords for a shape dict."""
    kw = {k: int(shape[k]) for k in ("block", "chunk", "walk_cap", "refill_min", "waves") if k in shape}
    kw["flags"] = (0 if shape.get("tile", True) else A.FLAG_NO_LDS_TILE) | \
        {"auto": 0, "direct": A.FLAG_DIRECT_TALLY, "binned": A.FLAG_BINNED_TALLY}[shape.get("tally", "auto")]
    if int(shape.get("copies", 1)) != 1:  # 1 per XCD is the library's default for one set
        kw["flags"] |= A.flag_grid_copies(int(shape["copies"]))
    return kw


def coordinate_descent(cost: Callable[[Dict], float], space: Dict[str, Sequence],
                       start: Dict, sweeps: int = 2, min_gain: float = 0.0
                       ) -> Tuple[Dict, List[Tuple[Dict, float]]]:
    """Minimise `cost` over the product space one knob at a time, `sweeps` passes over the
    knobs; each point is measured once (memoised).  A candidate replaces the best point only if
    it is cheaper by more than the fraction `min_gain` (timing noise does not move the shape).
    Returns (best shape, [(shape, cost)])."""
    memo: Dict[Tuple, float] = {}
    table: List[Tuple[Dict, float]] = []

    def f(sh: Dict) -> float:
        key = tuple(sorted(sh.items()))
        if key not in memo:
            memo[key] = float(cost(dict(sh)))
            table.append((dict(sh), memo[key]))
        return memo[key]

    best = dict(start)
    best_c = f(best)
    for _ in range(sweeps):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cand = dict(best)
                cand[knob] = v
                c = f(cand)
                if c < best_c * (1.0 - min_gain):
                    best, best_c, changed = cand, c, True
        if not changed:
            break
    return best, table


def fingerprint(scene) -> str:
    """Hash of everything the packet loop reads from the scene description."""
    d = scene.desc
    h = hashlib.sha1()

    def put(ptr, n, ctype):
        if n and ptr:
            h.update(C.string_at(ptr, n * C.sizeof(ctype)))

    put(d.nodes, d.n_nodes, A.Node)
    put(d.regions, d.n_regions, A.Region)
    put(d.materials, d.n_materials, A.Material)
    h.update(bytes(d.grid))
    h.update(bytes(d.source))
    return h.hexdigest()[:20]


def default_cache_path() -> str:
    base = os.environ.get("SMCRT_CACHE_DIR") or os.path.join(os.path.expanduser("~"), ".cache", "smcrt")
    return os.path.join(base, "autotune.json")


@dataclass
class TuneResult:
    shape: Dict
    kernel_ms: float           # best-of-reps kernel time of the probe at `shape`
    default_ms: float          # same at the library defaults
    table: List[Tuple[Dict, float]] = field(default_factory=list)
    cached: bool = False
    adjoint_ms: Dict[int, float] = field(default_factory=dict)  # adjoint probe time per waves

    @property
    def speedup(self) -> float:
        return self.default_ms / self.kernel_ms if self.kernel_ms > 0 else 1.0


def _device_name() -> str:
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}:{getattr(p, 'gcnArchName', '')}:{p.multi_processor_count}"


def tune(engine, n_probe: Optional[int] = None, reps: int = 2, sweeps: int = 2,
         cache_path: Optional[str] = None, use_cache: bool = True,
         space: Optional[Dict[str, Sequence]] = None, seed: int = 7919,
         target_ms: float = 150.0, min_gain: float = 0.02, adjoint: bool = True,
         adjoint_materials: Optional[Sequence[int]] = None) -> TuneResult:
    """Tune the launch shape of `engine`'s scene on the current GPU.

    n_probe: packets per probe (default: enough for `target_ms` of kernel time at the default
    shape, estimated from a first probe).  A persistent launch fills and drains: for the last
    few milliseconds of every launch most lanes are idle, which favours small blocks and chunks
    in short probes but not in long runs, so the probes are long enough for the steady state to
    dominate.  A knob moves away from the library default only for a gain above `min_gain`.
    The probe tallies into a scratch device grid.  With `adjoint`, the adjoint kernel's resident
    waves are tuned as well, on the forward's best shape (its register budget differs: it keeps
    per-material scores), and returned as the shape's "adj_waves" (used by Engine.grad), with
    the materials `adjoint_materials` differentiated (default every material, or the first 16:
    an adjoint launch tracks at most 16)."""
    import torch

    space = dict(space or SPACE)
    cache_path = cache_path or default_cache_path()
    key = f"{fingerprint(engine.scene)}|{_device_name()}|abi{A.ABI_VERSION}"
    if use_cache and os.path.exists(cache_path):
        try:
            entry = json.load(open(cache_path)).get(key)
        except (OSError, ValueError):
            entry = None
        if entry:
            return TuneResult(entry["shape"], entry["kernel_ms"], entry["default_ms"], [], cached=True)

    g = engine.scene.grid
    grid = torch.zeros(g.nxg * g.nyg * g.nzg, dtype=torch.float64, device="cuda")
    ref: Dict = {}

    def probe(shape: Dict, n: int) -> float:
        best = float("inf")
        for _ in range(reps):
            st = engine.run_device(n, grid.data_ptr(), seed=seed, first_packet=1 << 50,
                                   **run_kwargs(shape))
            inv = {k: st[k] for k in INVARIANT}
            if n in ref and ref[n] != inv:
                raise RuntimeError(f"launch shape {shape} changed the transported work: "
                                   f"{inv} != {ref[n]}")
            ref.setdefault(n, inv)
            best = min(best, st["kernel_ms"])
        return best

    start = {k: DEFAULT[k] for k in space}
    if n_probe is None:
        n = 200_000
        probe(start, n)  # warm-up: the LDS tile's pilot run, code load
        # calibrate at growing sizes until the size for target_ms is stable: short launches are
        # fill/drain bound, and large ones run the scene's run-time compiled kernels (faster than
        # the library's), so the rate measured at small sizes underestimates the probe's
        for _ in range(5):
            ms = probe(start, n)
            n_next = int(min(max(n * target_ms / max(ms, 1e-3), 200_000), 50_000_000))
            if abs(n_next - n) <= 0.15 * n:
                break
            n = min(n_next, 16 * n)
        n_probe = n_next
    t0 = time.perf_counter()
    best, table = coordinate_descent(lambda sh: probe(sh, n_probe), space, start, sweeps, min_gain)
    best_ms = dict((tuple(sorted(s.items())), c) for s, c in table)[tuple(sorted(best.items()))]
    if adjoint and "waves" in space:
        # same packets at every point: the counters must agree across the adjoint's probes
        adj: Dict[int, float] = {}
        adj_ref: Dict = {}
        mats = adjoint_materials
        if mats is None and engine.scene.n_materials > 16:
            mats = list(range(16))
        for w in space["waves"]:
            kw = run_kwargs(dict(best, waves=w))
            kw.pop("flags")  # the LDS tile option: the adjoint does not tally
            t = float("inf")
            for _ in range(reps):
                st = engine.grad(n_probe, seed=seed, first_packet=1 << 50, materials=mats, **kw).stats
                inv = {k: st[k] for k in INVARIANT}
                if adj_ref and adj_ref != inv:
                    raise RuntimeError(f"adjoint at {w} waves changed the transported work: {inv} != {adj_ref}")
                adj_ref = adj_ref or inv
                t = min(t, st["kernel_ms"])
            adj[w] = t
        fastest = min(adj, key=adj.get)
        aw = fastest if adj[fastest] < adj[best["waves"]] * (1.0 - min_gain) else best["waves"]
        best = dict(best, adj_waves=int(aw))
    default_ms = table[0][1]
    res = TuneResult(best, best_ms, default_ms, table, adjoint_ms=adj if adjoint and "waves" in space else {})
    if use_cache:
        try:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
            data = json.load(open(cache_path)) if os.path.exists(cache_path) else {}
        except (OSError, ValueError):
            data = {}
        data[key] = {"shape": best, "kernel_ms": best_ms, "default_ms": default_ms,
                     "n_probe": n_probe, "seconds": time.perf_counter() - t0}
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, cache_path)
    return res
