"""torch.autograd binding of the perturbation-MC adjoint (no reference counterpart).

    phi, r_d = MCRTFluence.apply(mua, mus, engine, n_packets, seed)

`mua`, `mus` are float64 tensors with one entry per material of `engine.scene` (the distinct
`mono` objects, in scene order).  The forward pass runs the packet loop with those properties and
returns `phi`, the raw path-length fluence per launched packet (shape (nx, ny, nz), FP64), and
`r_d`, the escaped-against-the-beam fraction.  The backward pass runs the adjoint kernel once
with voxel weights w = dL/dphi, which yields dL/dmu for every material in one launch:

    dL/dmu_r = sum_v w_v dphi_v/dmu_r + (dL/dr_d) dr_d/dmu_r

(score-function estimator; see signedmcrt_amd.engine.Engine.grad).  The backward launch reuses
the forward's random streams (same seed and packet ids), so forward and backward see the same
packets.
"""
from __future__ import annotations

import numpy as np
import torch


class MCRTFluence(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mua, mus, engine, n_packets: int, seed: int):
        mua_np = mua.detach().cpu().numpy().astype(np.float64)
        mus_np = mus.detach().cpu().numpy().astype(np.float64)
        engine.set_materials(mus=mus_np, mua=mua_np)
        r = engine.run(n_packets, seed=seed)
        ctx.engine, ctx.n, ctx.seed = engine, n_packets, seed
        ctx.shape = r.jmean.shape
        phi = torch.from_numpy(r.jmean / n_packets)
        rd = torch.tensor(r.stats["escaped_r"] / n_packets, dtype=torch.float64)
        return phi, rd

    @staticmethod
    def backward(ctx, grad_phi, grad_rd):
        eng = ctx.engine
        if grad_phi is None:
            w = np.zeros(ctx.shape, dtype=np.float32)
        else:
            w = grad_phi.detach().cpu().numpy().astype(np.float32)
        g = eng.grad(ctx.n, seed=ctx.seed, weights=w)
        grd = 0.0 if grad_rd is None else float(grad_rd)
        gmua = torch.from_numpy(g.d_w_dmua + grd * g.d_r_dmua)
        gmus = torch.from_numpy(g.d_w_dmus + grd * g.d_r_dmus)
        return gmua, gmus, None, None, None
