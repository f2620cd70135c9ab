"""Algorithmic work of the packet kernel, from its own event counters.

FLOPs per event are op counts of the kernel source (csrc/smcrt_device.h, csrc/packet_kernel.h):
FMA = 2, add/sub/mul/min/max = 1, each sqrt/rsqrt/rcp/log/sin/cos = 1 (integer work — the
Philox rounds, voxel indices, control flow — is not counted).  The counters are the kernel's:
sweeps (every region evaluated once), steps (march steps, each one DDA walk), crossings (voxel
chords tallied), scatters, interfaces, normals (gradient evaluations of one region), extra
region evaluations and launches.

Tally bytes: one FP64 atomic add per chord, a read-modify-write of 8 bytes (16 bytes of HBM
traffic when the line is not cached).
"""
from __future__ import annotations

from . import _abi as A

# FP32 FLOPs of one primitive evaluation, excluding its transform
PRIM_FLOPS = {
    A.SPHERE: 7,      # dot (1 mul + 2 fma = 5), sqrt, sub
    A.BOX: 16,        # 3 sub, 3 max0, length (5 + sqrt), 2 max, min0, add
    A.TORUS: 14,      # 2x (fma chain 3 + sqrt) + 2 sub + 2 mul
    A.CYLINDER: 46,   # pa (3), paba (5), w (6 fma = 12), |w| (6), x, y (4), x2/y2 (3), branch (3), sqrt, scale, sign
    A.TRIPRISM: 9,    # 2 fma, 4 max/sub, 3 abs-free ops
    A.SEGMENT: 23,    # pa (3), dot (5), scale, clamp (2), w (3 fma = 6), length (6)
    A.CAPSULE: 23,
    A.CONE: 52,
    A.EGG: 14,
    A.PLANE: 5,
    A.POLYGON: 16,    # grid test (4 cmp), cell index (2 fma), far-cell bound (2 sub, length 4, sub):
                      # the far-cell path only, so polygon work is a lower bound (near cells add
                      # ~15 per candidate edge, a count the event counters do not see)
}
TRANSFORM_FLOPS = 18                 # 3 components x 3 fma (identity transforms are skipped)
CSG_FLOPS = {A.UNION: 1, A.SMOOTH_UNION: 10, A.SUBTRACTION: 1, A.INTERSECTION: 1}
MOD_FLOPS = {A.EXTRUDE: 10, A.REVOLUTION: 7, A.ONION: 2, A.ELONGATE: 12, A.TWIST: 16, A.BEND: 16}
SWEEP_PER_REGION = 5                 # |d|, min |d|, min d, emission candidate (2)
STEP_FLOPS = 44                      # t = d*kappa, tau test, taurun, 3 fma position update, DDA set-up
                                     # (3 fma grid coords, 3 floor, 3 abs*scale, 3 rcp, 3x3 tMax)
CROSS_FLOPS = 6                      # min of 3 tMax, min with d, chord length, tMax advance, compare
SCATTER_FLOPS = 66                   # albedo test, HG inversion (rcp), clamp, sint, cos/sin, rotation,
                                     # renormalisation (rsqrt), new tau (log)
IFACE_FLOPS = 8                      # crossing step d + 2 eps and 3 fma
FRESNEL_FLOPS = 52                   # normalise gradient (rsqrt), Fresnel (2 sqrt, 2 div), refract
EMIT_FLOPS = 12                      # circular source: sqrt, sin/cos, 3x3 transform; plus first tau
DUAL_FACTOR = 4                      # forward-mode gradient: value + 3 partials


def node_flops(n: A.Node) -> int:
    if n.op == A.OP_PRIM:
        ident = list(n.t) == [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
        return PRIM_FLOPS[n.kind] + (0 if ident else TRANSFORM_FLOPS)
    if n.op == A.OP_CSG:
        return CSG_FLOPS[n.kind]
    if n.op in (A.OP_MOD_BEGIN, A.OP_MOD_END):
        return MOD_FLOPS[n.kind] // 2
    return 0


def region_flops(scene) -> list:
    out = []
    for r in scene._regions:
        out.append(sum(node_flops(scene._nodes[r.first_node + i]) for i in range(r.n_nodes)))
    return out


def sweep_flops(scene) -> int:
    return sum(f + SWEEP_PER_REGION for f in region_flops(scene))


def flops(scene, st: dict) -> float:
    """Algorithmic FP32 FLOPs of a run with counters `st`."""
    rf = region_flops(scene)
    avg_region = sum(rf) / len(rf)
    return (st["ev_sweeps"] * sweep_flops(scene)
            + st["ev_steps"] * STEP_FLOPS
            + st["ev_crossings"] * CROSS_FLOPS
            + st["scatters"] * SCATTER_FLOPS
            + st["ev_interfaces"] * IFACE_FLOPS
            + st["ev_normals"] * (DUAL_FACTOR * avg_region + FRESNEL_FLOPS)
            + st["ev_region_evals"] * avg_region
            + st["launched"] * EMIT_FLOPS)


def tally_bytes(st: dict) -> float:
    """Bytes moved by the fluence tally: one 8-byte FP64 read-modify-write per chord."""
    return st["ev_crossings"] * 16.0
