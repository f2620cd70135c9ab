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


def per_packet_rel_sd(batches):
    """Median over voxels with fluence >= 1% of max of the per-packet relative sd, from
    equally sized batch tallies (list of (grid, n_packets))."""
    import numpy as np
    a = np.stack([g / n for g, n in batches])
    m = a.mean(axis=0)
    mask = m >= 0.01 * m.max()
    sd = a[:, mask].std(axis=0, ddof=1) / m[mask]
    n_b = batches[0][1]
    return float(np.median(sd) * np.sqrt(n_b))


def kernel_label(eng, waves):
    """The forward kernel's name as rocprof reports it: the library's variant, or the kernel
    compiled at run time for the scene's program (instantiated on the scene's exact kinds)."""
    name, scene_mask, variant_mask = eng.kernel_variant()
    if "compiled" in name:
        return "packet_kernel<false, %du, %d, SceneSweep> (%s)" % (scene_mask, waves, name)
    return "packet_kernel<false, %du, %d> (%s)" % (variant_mask, min(waves, 3), name)


def read_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", f"pmc_packet_kernel_{workload}.md")
    if not os.path.exists(p):
        return None
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read())
    return float(m.group(1)) if m else None


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
    ap.add_argument("--error-batches", type=int, default=10)
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
        res = eng.tune(use_cache=False)
        tuned = {"shape": res.shape, "probe_kernel_ms": res.kernel_ms,
                 "default_probe_kernel_ms": res.default_ms, "speedup_vs_default": res.speedup,
                 "points": len(res.table), "seconds": time.perf_counter() - t_tune}

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
    traffic = read_traffic(args.workload)

    # per-packet fluence error of the GPU estimator: B independent batches
    B = args.error_batches
    sd_gpu = None
    if B >= 2:
        nb = max(n // B, 100_000)
        gb = []
        for b in range(B):
            gg = torch.zeros(nv, dtype=torch.float64, device="cuda")
            eng.run_device(nb, gg.data_ptr(), seed=99, first_packet=10**9 + b * nb)
            gb.append((gg.cpu().numpy(), nb))
        sd_gpu = per_packet_rel_sd(gb)
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
        meas = np.zeros(nv)
        eng.run_device(nf, d_grid.zero_().data_ptr(), seed=4242)
        meas = d_grid.cpu().numpy().reshape(g.shape, order="F") / nf
        fscene, _ = scenes.build(args.workload, mua=[0.5] * 16)
        feng = Engine(fscene, shape=eng.shape)
        t_fit = time.perf_counter()
        hist = fit_mua(feng, meas, materials=list(range(16)), mua0=0.5, n_packets=nf,
                       iters=args.fit_iters, lr=0.1, block=10)
        err = np.abs(hist.final / truth - 1)
        fit = {"iters": args.fit_iters, "packets_per_launch": nf, "seconds": time.perf_counter() - t_fit,
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
        sd_cpu = per_packet_rel_sd(cb)
        cpu_rate = nbc * B / secs
        cpu = {"value": cpu_rate, "unit": "packets/s", "cores": threads, "kind": "port",
               "sample": f"{nbc * B} packets of the same workload in {B} batches, FP64 C "
                         f"restatement of the reference loop (oracle/), OpenMP on {threads} "
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
                     "traffic": traffic, "kernel": kernel_label(eng, int((tuned or {}).get("shape", {}).get("waves", 3))),
                     "kernel_ms": kern_ms, "flops_per_launch": flops_launch},
        "roofline_tally": {"bound": "hbm", "achieved": tally_gbs, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": tally_gbs / HBM_PEAK_GBS,
                           "bytes_per_launch": tally_launch,
                           "crossings_per_launch": tot["ev_crossings"] / args.steps},
        "events_per_packet": {k: tot[k] / tot["launched"] for k in
                              ("scatters", "ev_sweeps", "ev_steps", "ev_crossings",
                               "ev_interfaces", "ev_normals")},
        "fractions": {k: tot[k] / tot["launched"] for k in
                      ("absorbed", "escaped_r", "escaped_t", "killed_surface")},
        "per_packet_rel_sd": sd_gpu,
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
        res = engines[1].tune(use_cache=False)
        shape = {k: v for k, v in res.shape.items() if k != "tile"}  # batches tally straight to HBM
        tuned = {"shape": res.shape, "tuned_on": "refracting_sphere 64^3", "speedup_vs_default": res.speedup}
    b = Batch(engines, sets, seed=123456789, **shape)
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
    out = {
        "metric": METRIC, "value": value, "unit": "packets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched", "sets": len(sets), "packets_per_step": int(total),
                   "budget_min": int(budgets.min()), "budget_max": int(budgets.max()),
                   "grids": "64^3 FP64 per set", "geometries": ["slab", "refracting_sphere"],
                   "launch": "one HIP graph: fill kernels + one persistent packet-kernel node per geometry",
                   "launch_shape": shape or "library default"},
        "fractions": frac, "autotune": tuned, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
