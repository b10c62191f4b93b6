"""torch.autograd.Function around the adjoint entry point (fq_backward).

    E = qaoa_expectation(gammas, betas, graph, dtype="c64")
    E.backward()          # gammas.grad, betas.grad from the reverse uncompute

The angle tensors are small; they are read on the host (CPU tensors, or
detached copies of device tensors) and the state vector lives in the
library's own HBM allocation for the lifetime of the autograd node.  The
forward pass is fq_run(..., FQ_SCHED_Z2) (FQ_SCHED_FUSED for N < 3) +
fq_expect_cut; the backward
pass is fq_backward, which walks the state back layer by layer (no
checkpoints).  The reference has no expectation or gradient path; this is
the differentiation form named for this build.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Graph, State


class QAOAExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gammas, betas, graph, dtype, C):
        g = gammas.detach().to("cpu", torch.float64).numpy().copy()
        b = betas.detach().to("cpu", torch.float64).numpy().copy()
        if g.shape != b.shape or g.ndim != 1:
            raise ValueError("gammas and betas must be 1-D of equal length")
        st = State(graph.N, C=C if C is not None else (13 if dtype == "c128" else 14), dtype=dtype)
        # the half-state schedule leaves the same identity-layout final state
        st.run(graph, g, b, "z2" if graph.N >= 3 else "fused")
        e, _ = st.expect_cut(graph)
        ctx.st = st
        ctx.graph = graph
        ctx.angles = (g, b)
        ctx.devices = (gammas.device, betas.device)
        return torch.tensor(e, dtype=torch.float64, device=gammas.device)

    @staticmethod
    def backward(ctx, grad_out):
        g, b = ctx.angles
        st = ctx.st
        ctx.st = None
        try:
            _, gg, gb = st.backward(ctx.graph, g, b)
        finally:
            st.destroy()
        scale = float(grad_out.detach().to("cpu"))
        dg = torch.from_numpy(np.asarray(gg) * scale).to(ctx.devices[0])
        db = torch.from_numpy(np.asarray(gb) * scale).to(ctx.devices[1])
        return dg, db, None, None, None


def qaoa_expectation(gammas, betas, graph: Graph, dtype="c64", C=None):
    """<C>(gammas, betas) of weighted Max-Cut QAOA on `graph`, differentiable
    in the angles through the adjoint kernel."""
    return QAOAExpectation.apply(gammas, betas, graph, dtype, C)
