"""signedmcrt_amd — MI355X-native photon-packet engine for signedMCRT's packet loop.

The compute path is libsmcrt.so (hand-written HIP kernels for gfx950 behind the C ABI of
include/smcrt.h).  This package holds the host-side mirror of the reference's interfaces: the SDF
model API (sdf), scene assembly (scene, scenes), the TOML input and the fluence/detector output
files it drives (config), the run driver (engine), detectors (detectors), the adjoint's autograd
binding (autograd), the inverse fit (inverse), the batched launcher (batch), the autotuner
(autotune) and the roofline accounting (roofline).
"""
from .sdf import (box, bend, capsule, cone, cylinder, egg, elongate, extrude, identity,
                  intersection, invert, model, mono, onion, plane, revolution, rotate_x, rotate_y,
                  rotate_z, rotationAlign, rotmat, segment, skewSymm, SmoothUnion, sphere,
                  subtraction, torus, translate, triprism, twist, union)
from .scene import Scene, cart_grid, source

__all__ = ["box", "bend", "capsule", "cone", "cylinder", "egg", "elongate", "extrude", "identity",
           "intersection", "invert", "model", "mono", "onion", "plane", "revolution", "rotate_x",
           "rotate_y", "rotate_z", "rotationAlign", "rotmat", "segment", "skewSymm",
           "SmoothUnion", "sphere", "subtraction", "torus", "translate", "triprism", "twist",
           "union", "Scene", "cart_grid", "source"]
