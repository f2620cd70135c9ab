#!/usr/bin/env python3
"""Throughput of the MI355X packet loop on the skin + vessels workload, and the adjoint's cost.

A step is one persistent launch of the packet kernel over 1e8 packets of the skin + vessels scene
(five stacked tissue boxes plus eight blood-vessel cylinders merged by SmoothUnion, in air;
pencil beam), tallying the path-length fluence into a device-resident 256^3 FP64 grid.  Inputs
(scene program, material table, grid) are resident in HBM before timing starts; the launch shape
comes from the project's autotuner (tuned before timing, reported).

Prints one JSON line: packets/s, the roofline of the packet kernel (FP32 VALU rate from the
kernel's own event counters against the MI355X FP32 peak, kernel time from HIP events on the
kernel's stream), the FP64 tally's byte rate against HBM bandwidth, the CPU baseline (the FP64 C
restatement of the reference loop, oracle/, timed on this host's cores on a bounded sample of the
same workload, with the per-packet fluence error of both sides so throughput is compared at equal
fluence error), and the gradient cost: one adjoint launch giving d(R_d, fluence sum)/d(mu_a, mu_s)
of every material against the 2P forward launches of central finite differences.

    python bench.py [--steps K] [--warmup W] [--workload slab|refracting_sphere|skin_vessels]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 (published)
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E (published)
METRIC = "photon packets/s @ fixed fluence error (1 GPU vs host CPU); grad cost vs FD"
# workload -> (packets per step, description, gradient: (materials the adjoint differentiates,
# number of parameters central finite differences would perturb) or None)
WORKLOADS = {
    "slab": (1_000_000, "slab 1x1x0.2 cm box mus=100 mua=0.1 g=0.9 n=1.37 in air, pencil beam, "
             "100^3 grid", None),
    "refracting_sphere": (10_000_000, "clear sphere r=0.5 n=1.5 mua=0.01 in 2 cm cube mus=50 "
                          "mua=0.05 g=0.8 n=1.33, air, circular beam r=0.3 cm, 200^3 grid", None),
    "skin_vessels": (100_000_000, "5 stacked tissue boxes (0.01-0.5 cm, mua 0.02-5, mus 20-300, "
                     "g 0.8-0.95, n 1.37-1.45, seeded) + 8 blood cylinders r=0.01-0.05 mua=20 "
                     "merged by SmoothUnion k=0.01, air, pencil beam, 256^3 grid",
                     (None, 10)),  # every material; FD on the 5 layers' (mu_a, mu_s)
    "polygon_regions": (100_000_000, "16 seeded closed polygons x 125 edges (2000 segments) "
                        "extruded to 0.5 cm in a 4x4x0.5 cm tissue slab (16 materials + background), "
                        "uniform-grid polygon accelerator, uniform top illumination, 400x400x50 grid",
                        (list(range(16)), 16)),  # the 16 polygons' mu_a
}


# Voxels the per-packet fluence error is taken over: fluence >= this fraction of the maximum.  A
# pencil beam's fluence falls four decades within a few voxels of its axis (1 % of the maximum
# keeps 150-470 voxels of skin + vessels and the slab, too few for a stable median), so the floor
# is set per workload to keep >= 1e4 voxels (the same floors as the per-voxel parity tests).
ERROR_FLOOR = {"slab": 1e-3, "refracting_sphere": 1e-2, "skin_vessels": 3e-4, "polygon_regions": 1e-2}


def per_packet_rel_sd(batches, floor):
    """Median over voxels with fluence >= floor x max of the per-packet relative sd, from
    equally sized batch tallies (list of (grid, n_packets)); also the number of voxels."""
    import numpy as np
    s = s2 = None
    for g, n in batches:
        x = g / n
        s = x.copy() if s is None else s + x
        s2 = x * x if s2 is None else s2 + x * x
    k = len(batches)
    m = s / k
    var = np.maximum(s2 / k - m * m, 0.0) * k / (k - 1)
    mask = m >= floor * m.max()
    sd = np.sqrt(var[mask]) / m[mask]
    n_b = batches[0][1]
    return float(np.median(sd) * np.sqrt(n_b)), int(mask.sum())


def atomic_ceiling(nv, n=1 << 26, reps=3):
    """Measured FP64 global-atomic ceiling: n adds at uniformly random addresses of an nv-double
    grid (torch's index_add_, one atomic per element), best of `reps`, in atomics/s."""
    import torch
    idx = torch.randint(0, nv, (n,), device="cuda", dtype=torch.int64)
    src = torch.ones(n, device="cuda", dtype=torch.float64)
    g = torch.zeros(nv, device="cuda", dtype=torch.float64)
    g.index_add_(0, idx, src)  # warm
    best = float("inf")
    for _ in range(reps):
        g.zero_()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.index_add_(0, idx, src)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    assert g.sum().item() == n
    del idx, src, g
    return n / (best * 1e-3)


JIT_MIN_PACKETS = 1 << 19  # the library's threshold for its run-time compiled kernels


def kernel_label(eng, waves, n_packets):
    """The forward kernel a launch of n_packets ran, named as rocprof reports it: the library's
    variant, or the kernel compiled at run time for the scene's program (instantiated on the
    scene's exact kinds), which launches of >= JIT_MIN_PACKETS use once it is built (SMCRT_JIT=1:
    every launch, =0: none)."""
    name, scene_mask, variant_mask = eng.kernel_variant()
    jit = os.environ.get("SMCRT_JIT", "")
    if "compiled" in name and "failed" not in name and (jit == "1" or (jit != "0" and n_packets >= JIT_MIN_PACKETS)):
        # the last BIN parameter: the binned tally's stage (the library reports the last launch's)
        return "packet_kernel<false, %du, %d, SceneSweep, %s> (%s)" % (
            scene_mask, waves, "true" if "+binned" in name else "false", name)
    return "packet_kernel<false, %du, %d, InterpretedSweep, false> (%s)" % (variant_mask, min(waves, 3), name)


def batch_kernel_label(eng, waves, n_packets):
    """The kernel a batch's geometry group of n_packets ran: the scene's run-time compiled kernel
    for several sets (groups of >= JIT_MIN_PACKETS, once built; SMCRT_JIT=1: every group, =0: none)
    or the library's variant."""
    name, scene_mask, variant_mask = eng.kernel_variant()
    jit = os.environ.get("SMCRT_JIT", "")
    if "compiled" in name and "failed" not in name and (jit == "1" or (jit != "0" and n_packets >= JIT_MIN_PACKETS)):
        return "packet_kernel<false, %du, %d, SceneSweepSets, false>" % (scene_mask, waves)
    return "packet_kernel<false, %du, %d, InterpretedSweep, false>" % (variant_mask, min(waves, 3))


def read_traffic(workload, label):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this workload, when the
    summary's kernel is the one this run timed (its `kernel:` line equals `label` up to the
    variant note in parentheses); None otherwise."""
    p = os.path.join(ROOT, "profiles", f"pmc_packet_kernel_{workload}.md")
    if not os.path.exists(p):
        return None
    text = open(p).read()
    k = re.search(r"^kernel:\s*(.+?)\s*$", text, re.M)
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", text)
    if not (k and m) or k.group(1).split(" (")[0] != label.split(" (")[0]:
        return None
    return float(m.group(1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--packets", type=int, default=0, help="packets per step (0: the workload's)")
    ap.add_argument("--workload", default="skin_vessels", choices=sorted(WORKLOADS) + ["batched"])
    ap.add_argument("--sets", type=int, default=256, help="batched: property sets")
    ap.add_argument("--fit-iters", type=int, default=0,
                    help="polygon_regions: inverse-fit iterations to report (0: skip)")
    ap.add_argument("--fit-packets", type=int, default=20_000_000)
    ap.add_argument("--no-tune", action="store_true", help="library default launch shape")
    ap.add_argument("--shape", default="", help="launch shape as JSON (skips the autotuner), e.g. "
                    "'{\"block\": 256, \"chunk\": 256, \"walk_cap\": 8, \"refill_min\": 1, \"tile\": true}'")
    ap.add_argument("--grad-steps", type=int, default=2,
                    help="adjoint launches over the first timed step's packets (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-packets", type=int, default=2_000_000)
    ap.add_argument("--error-batches", type=int, default=20)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB device copy (PMC byte-counter calibration) and exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py measures one GPU: this engine runs on a single MI355X (asked for "
              f"{max(args.gpus, world)}); multi-GPU sharding is a design note only.", flush=True)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        n = 1 << 27  # 1 GiB of float64
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return

    from signedmcrt_amd import roofline, scenes
    from signedmcrt_amd.engine import Engine

    if args.workload == "batched":
        return bench_batched(args)

    n_default, description, grad_spec = WORKLOADS[args.workload]
    scene, _ = scenes.build(args.workload)
    g = scene.grid
    nv = g.nxg * g.nyg * g.nzg
    eng = Engine(scene)
    d_grid = torch.zeros(nv, dtype=torch.float64, device="cuda")
    n = args.packets or n_default

    tuned = None
    if args.shape:
        eng.shape = json.loads(args.shape)
        tuned = {"shape": eng.shape, "given": True}
    elif not args.no_tune:
        t_tune = time.perf_counter()
        res = eng.tune(use_cache=False, adjoint=bool(grad_spec and args.grad_steps > 0),
                       adjoint_materials=grad_spec[0] if grad_spec else None)
        tuned = {"shape": res.shape, "probe_kernel_ms": res.kernel_ms,
                 "default_probe_kernel_ms": res.default_ms, "speedup_vs_default": res.speedup,
                 "points": len(res.table), "seconds": time.perf_counter() - t_tune}
        if res.adjoint_ms:
            tuned["adjoint_probe_kernel_ms"] = {str(w): t for w, t in sorted(res.adjoint_ms.items())}

    for w in range(args.warmup):
        eng.run_device(n, d_grid.data_ptr(), seed=1000 + w, first_packet=0)
    torch.cuda.synchronize()
    d_grid.zero_()
    torch.cuda.synchronize()

    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(eng.run_device(n, d_grid.data_ptr(), seed=123456789, first_packet=k * n))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    value = args.steps * n / wall
    kern_ms = float(np.mean([s["kernel_ms"] for s in stats]))
    tot = {k: sum(s[k] for s in stats) for k in stats[0] if not isinstance(stats[0][k], list)}

    flops_launch = roofline.flops(scene, tot) / args.steps
    achieved_tf = flops_launch / (kern_ms * 1e-3) / 1e12
    tally_launch = roofline.tally_bytes(tot) / args.steps
    tally_gbs = tally_launch / (kern_ms * 1e-3) / 1e9
    label = kernel_label(eng, int((tuned or {}).get("shape", {}).get("waves", 3)), n)
    traffic = read_traffic(args.workload, label)
    ceiling = atomic_ceiling(nv)
    atomics_launch = tot["ev_atomics"] / args.steps

    # per-packet fluence error of the GPU estimator: B independent batches
    B = args.error_batches
    floor = ERROR_FLOOR[args.workload]
    sd_gpu = None
    if B >= 2:
        nb = max(n // B, 100_000)
        gb = []
        for b in range(B):
            gg = torch.zeros(nv, dtype=torch.float64, device="cuda")
            eng.run_device(nb, gg.data_ptr(), seed=99, first_packet=10**9 + b * nb)
            gb.append((gg.cpu().numpy(), nb))
        sd_gpu, nvox_err = per_packet_rel_sd(gb, floor)
        del gb

    # gradient cost: the adjoint launch over exactly the packets of the first timed step (same
    # seed and ids: the same histories) against that step's forward launch.  Central finite
    # differences need 2 forward launches per parameter.  (Comparing equal packet sets matters:
    # a launch lasts at least as long as its longest history, ~1e5 scatters in a low-absorption
    # layer, which for small launches can dominate the kernel time.)
    grad = None
    if grad_spec and args.grad_steps > 0:
        gmats, p = grad_spec
        ng = n
        fwd_ms = stats[0]["kernel_ms"]
        eng.grad(min(ng, 1_000_000), seed=77, materials=gmats)  # warm (code load, LDS sizing)
        gr = [eng.grad(ng, seed=123456789, first_packet=0, materials=gmats) for _ in range(args.grad_steps)]
        grad_ms = float(np.mean([x.stats["kernel_ms"] for x in gr]))
        for k in ("launched", "absorbed", "escaped_r", "scatters"):
            assert gr[0].stats[k] == stats[0][k], ("adjoint and forward histories differ", k)
        nsel = scene.n_materials if gmats is None else len(gmats)
        grad = {"packets": ng, "forward_kernel_ms": fwd_ms, "grad_kernel_ms": grad_ms,
                "grad_cost_vs_forward": grad_ms / fwd_ms,
                "fd_parameters": p, "fd_forward_launches": 2 * p,
                "speedup_vs_central_fd": 2 * p * fwd_ms / grad_ms,
                "derivatives_per_launch": 4 * nsel,
                "d_rd_dmua": [float(v) for v in gr[0].d_r_dmua],
                "d_rd_dmus": [float(v) for v in gr[0].d_r_dmus]}

    fit = None
    if args.workload == "polygon_regions" and args.fit_iters > 0:
        from signedmcrt_amd.inverse import fit_mua
        truth = np.array([scene.materials[k].mua for k in range(16)])
        nf = args.fit_packets
        nm = 5 * nf  # the "measurement": the engine's own tally at the true mu_a, 5x the fit's packets
        eng.run_device(nm, d_grid.zero_().data_ptr(), seed=4242)
        meas = d_grid.cpu().numpy().reshape(g.shape, order="F") / nm
        fscene, _ = scenes.build(args.workload, mua=[0.5] * 16)
        feng = Engine(fscene, shape=eng.shape)
        t_fit = time.perf_counter()
        hist = fit_mua(feng, meas, materials=list(range(16)), mua0=0.5, n_packets=nf,
                       iters=args.fit_iters, lr=0.1, lr_final=0.02,
                       average_last=max(1, args.fit_iters // 5), block=10)
        err = np.abs(hist.final / truth - 1)
        fit = {"iters": args.fit_iters, "packets_per_launch": nf, "measurement_packets": nm,
               "optimizer": "Adam on log mu_a, lr 0.1 -> 0.02 (cosine), mean of the last 20 % of iterates",
               "seconds": time.perf_counter() - t_fit,
               "loss_first": hist.loss[0], "loss_last": hist.loss[-1],
               "rel_err_start": float(np.median(np.abs(0.5 / truth - 1))),
               "rel_err_median": float(np.median(err)), "rel_err_max": float(err.max()),
               "truth": truth.tolist(), "fit": hist.final.tolist()}

    cpu = None
    if not args.no_cpu and B >= 2:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        threads = len(os.sched_getaffinity(0))
        nc = args.cpu_packets
        nbc = nc // B
        cb, secs = [], 0.0
        for b in range(B):
            r = pyoracle.run(scene, nbc, seed=4321, first_packet=b * nbc, threads=threads)
            secs += r["seconds"]
            cb.append((r["jmean"].reshape(-1, order="F"), nbc))
        sd_cpu, _ = per_packet_rel_sd(cb, floor)
        cpu_rate = nbc * B / secs
        cpu = {"value": cpu_rate, "unit": "packets/s", "cores": threads, "kind": "port",
               "sample": f"{nbc * B} packets of the same workload in {B} batches, FP64 C "
                         f"restatement of the reference loop (oracle/; the CPU port, not the "
                         f"reference, which cannot be built here), OpenMP on {threads} "
                         f"host threads, packet loop only (omp_get_wtime)",
               "per_packet_rel_sd": sd_cpu,
               "speedup_packets": value / cpu_rate,
               "speedup_at_equal_error": (value / cpu_rate) * (sd_cpu / sd_gpu) ** 2}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "packets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "packets_per_step": n,
                   "grid": [g.nxg, g.nyg, g.nzg], "tally": "fp64", "scene": description,
                   "launch_shape": tuned["shape"] if tuned else "library default"},
        "roofline": {"bound": "valu", "achieved": achieved_tf, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tf / FP32_PEAK_TFLOPS,
                     "traffic": traffic, "kernel": label,
                     "kernel_ms": kern_ms, "flops_per_launch": flops_launch},
        "roofline_tally": {"bound": "hbm", "achieved": tally_gbs, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": tally_gbs / HBM_PEAK_GBS,
                           "bytes_per_launch": tally_launch,
                           "crossings_per_launch": tot["ev_crossings"] / args.steps},
        "roofline_atomics": {"bound": "fp64 global atomics", "achieved": atomics_launch / (kern_ms * 1e-3),
                             "peak": ceiling, "unit": "atomics/s",
                             "frac": atomics_launch / (kern_ms * 1e-3) / ceiling,
                             "atomics_per_launch": atomics_launch,
                             "peak_from": "measured in this run: 2^26 FP64 adds at random addresses of a "
                                          "grid of this workload's size (torch index_add_)"},
        "events_per_packet": {k: tot[k] / tot["launched"] for k in
                              ("scatters", "ev_sweeps", "ev_steps", "ev_crossings",
                               "ev_interfaces", "ev_normals", "ev_atomics")},
        "fractions": {k: tot[k] / tot["launched"] for k in
                      ("absorbed", "escaped_r", "escaped_t", "killed_surface")},
        "per_packet_rel_sd": sd_gpu,
        "error_voxels": {"floor": floor, "count": nvox_err if B >= 2 else None},
        "cpu_baseline": cpu,
        "grad": grad,
        "autotune": tuned,
        "inverse_fit": fit,
    }
    print(json.dumps(out), flush=True)


def batched_sets(n_sets: int, seed: int = 20240917):
    """The batched workload's property sets, drawn from a seeded generator: mu_a log-uniform
    0.01-1/cm, mu_s uniform 10-300/cm, g uniform 0.7-0.95, n uniform 1.33-1.45, packet budget
    log-uniform 1e5-1e7; even sets on the slab geometry, odd sets on the refracting sphere
    (the drawn properties go to the scattering medium: the slab, the sphere's cube)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_sets):
        mua = float(np.exp(rng.uniform(np.log(0.01), np.log(1.0))))
        mus = float(rng.uniform(10.0, 300.0))
        g = float(rng.uniform(0.7, 0.95))
        n = float(rng.uniform(1.33, 1.45))
        budget = int(np.exp(rng.uniform(np.log(1e5), np.log(1e7))))
        out.append((i % 2, mua, mus, g, n, budget))
    return out


def batch_shape(shape):
    """The launch-shape knobs of a tuned single-launch shape that a batch takes: block size and
    walk cap (batches tally straight to HBM: no LDS tile, no binned tally; no adjoint; their wave
    count, chunk and refill policy are their own, `Batch.fastest` and the library's batch
    defaults)."""
    return {k: v for k, v in shape.items() if k in ("block", "walk_cap")}


def bench_batched(args):
    """One HIP-graph replay = every set's packets in one persistent launch per geometry."""
    import numpy as np
    import torch
    from signedmcrt_amd import scenes
    from signedmcrt_amd.batch import Batch, BatchSet
    from signedmcrt_amd.engine import Engine

    engines = [Engine(scenes.build("slab", nxyz=64)[0]), Engine(scenes.build("refracting_sphere", nxyz=64)[0])]
    medium = [0, 1]  # index of the scattering medium among each scene's materials
    sets = []
    first = 0
    for geo, mua, mus, g, n, budget in batched_sets(args.sets):
        mats = engines[geo].scene.materials
        MUS = [m.mus for m in mats]
        MUA = [m.mua for m in mats]
        HG = [m.hgg for m in mats]
        NN = [m.n for m in mats]
        k = medium[geo]
        MUS[k], MUA[k], HG[k], NN[k] = mus, mua, g, n
        sets.append(BatchSet(geo, MUS, MUA, budget, first_packet=first, hgg=HG, n=NN))
        first += budget
    total = sum(s.n_packets for s in sets)
    shape = {}
    tuned = None
    if not args.no_tune:
        res = engines[1].tune(use_cache=False, adjoint=False)
        shape = batch_shape(res.shape)
        tuned = {"shape": res.shape, "tuned_on": "refracting_sphere 64^3", "speedup_vs_default": res.speedup}
    if args.no_tune:
        b = Batch(engines, sets, seed=123456789, **shape)
    else:
        # the batch's own wave count and refill policy (Batch.fastest's candidates)
        b = Batch.fastest(engines, sets, seed=123456789, **shape)
        tuned["batch_shape"] = b.shape
    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = total * args.steps / wall
    st = [b.read(i)[1] for i in range(len(sets))]
    launched = sum(x["launched"] for x in st)
    assert launched == total, (launched, total)
    frac = {k: sum(x[k] for x in st) / total for k in ("absorbed", "escaped_r", "escaped_t", "killed_surface")}
    budgets = np.array([s.n_packets for s in sets])

    # roofline of the replay: counted FLOPs of each geometry group's launch (its event counters,
    # its sets' scatters and packets) over the replay's device time (both packet-kernel nodes back
    # to back, plus the grid clears and the partial-grid reduction)
    from signedmcrt_amd import roofline
    flops = tally = 0.0
    for geo, eng in enumerate(engines):
        idx = [i for i, s in enumerate(sets) if s.scene == geo]
        if not idx:
            continue
        g = dict(st[idx[0]])  # ev_* and err_*: the group's
        for k in ("launched", "scatters"):
            g[k] = sum(st[i][k] for i in idx)
        assert g["err_unreachable"] == 0 and g["err_nan"] == 0
        flops += roofline.flops(eng.scene, g)
        tally += roofline.tally_bytes(g)
    replay_ms = st[0]["kernel_ms"]
    achieved = flops / (replay_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
            "kernel": " + ".join(batch_kernel_label(eng, int(b.shape.get("waves", 3)),
                                                    sum(s.n_packets for s in sets if s.scene == geo))
                                 for geo, eng in enumerate(engines)) + " (one graph replay)",
            "kernel_ms": replay_ms, "flops_per_launch": flops}
    roof_tally = {"bound": "hbm", "achieved": tally / (replay_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": tally / (replay_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                  "bytes_per_launch": tally}

    cpu = None
    if not args.no_cpu:
        # the oracle on every set at 1/200 of its budget (at least 200 packets): the same mix of
        # sets as the replay, weighted by budget
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import copy
        import pyoracle
        from signedmcrt_amd.scene import Scene
        threads = len(os.sched_getaffinity(0))
        secs, packets = 0.0, 0
        for i, s in enumerate(sets):
            sc = engines[s.scene].scene
            tmp = Scene(copy.deepcopy(sc.array), sc.grid, sc.source)
            for j, m in enumerate(tmp.materials):
                m.mus, m.mua, m.hgg, m.n = s.mus[j], s.mua[j], s.hgg[j], s.n[j]
            tmp = Scene(tmp.array, tmp.grid, tmp.source)
            nn = max(s.n_packets // 200, 200)
            r = pyoracle.run(tmp, nn, seed=4321, first_packet=s.first_packet, threads=threads)
            secs += r["seconds"]
            packets += nn
        cpu = {"value": packets / secs, "unit": "packets/s", "cores": threads, "kind": "port",
               "sample": f"{packets} packets: every set at 1/200 of its budget (>= 200), FP64 C "
                         f"restatement of the reference loop (oracle/; the CPU port, not the "
                         f"reference), OpenMP on {threads} host threads, packet loop only",
               "speedup_packets": value / (packets / secs)}
    out = {
        "metric": METRIC, "value": value, "unit": "packets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched", "sets": len(sets), "packets_per_step": int(total),
                   "budget_min": int(budgets.min()), "budget_max": int(budgets.max()),
                   "grids": "64^3 FP64 per set", "geometries": ["slab", "refracting_sphere"],
                   "launch": "one HIP graph: fill kernels + one persistent packet-kernel node per geometry",
                   "launch_shape": b.shape or "library default"},
        "fractions": frac, "autotune": tuned, "roofline": roof, "roofline_tally": roof_tally,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
