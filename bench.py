#!/usr/bin/env python3
"""Throughput of the MI355X packet loop on the refracting-sphere workload.

A step is one persistent launch of the packet kernel over 1e7 packets of the refracting-sphere
scene (clear sphere r = 0.5 cm, n = 1.5, in a 2 cm scattering cube, circular beam r = 0.3 cm),
tallying the path-length fluence into a device-resident 200^3 FP64 grid.  Inputs (scene
program, material table, grid) are resident in HBM before timing starts.

Prints one JSON line: packets/s, the roofline of the packet kernel (FP32 VALU rate from the
kernel's own event counters against the MI355X FP32 peak, kernel time from HIP events on the
kernel's stream), the FP64 tally's byte rate against HBM bandwidth, and the CPU baseline: the
FP64 C restatement of the reference loop (oracle/) timed on this host's cores on a bounded
sample of the same workload, with the per-packet fluence error of both sides so throughput is
compared at equal fluence error.

    python bench.py [--steps K] [--warmup W] [--no-cpu]
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
    ap.add_argument("--packets", type=int, default=10_000_000)
    ap.add_argument("--workload", default="refracting_sphere")
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

    scene, _ = scenes.build(args.workload)
    g = scene.grid
    nv = g.nxg * g.nyg * g.nzg
    eng = Engine(scene)
    d_grid = torch.zeros(nv, dtype=torch.float64, device="cuda")
    n = args.packets

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
                   "grid": [g.nxg, g.nyg, g.nzg], "tally": "fp64",
                   "scene": "clear sphere r=0.5 n=1.5 mua=0.01 in 2 cm cube mus=50 mua=0.05 "
                            "g=0.8 n=1.33, air, circular beam r=0.3 cm"},
        "roofline": {"bound": "valu", "achieved": achieved_tf, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved_tf / FP32_PEAK_TFLOPS,
                     "traffic": traffic, "kernel": "packet_kernel<false>",
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
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
