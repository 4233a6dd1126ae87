"""torch.autograd face of the discrete adjoint (fvh_objective / fvh_adjoint).

``flow_objective(solver, q0, n_steps, wx, wy, xy=None)`` returns the scalar
objective J = sum_b p(ghost_b) (wx_b n_x + wy_b n_y) len_b after n_steps RK
steps of the solver's run from the initial state q0, as a float64 tensor that
takes part in autograd:

- dJ/dq0 (one value per cell and conserved variable, reference layout) comes
  from the library's reverse-mode kernels with binomial checkpointing;
- when ``xy`` (the mesh's node coordinates, shape (n_nodes, 2)) is passed and
  requires grad, dJ/dxy comes from the same reverse sweep (geometry adjoint).
  The solver's mesh is fixed: ``xy`` must hold its node coordinates, and
  the derivative is taken at them.

When any input requires grad, the forward pass runs the adjoint straight
away (it evaluates J on the way) and keeps the gradients for backward; it
runs the plain objective otherwise.  Inputs may live on any device; states
travel to the host for the C ABI, which takes host buffers.
"""
from __future__ import annotations

import numpy as np
import torch


class FlowObjective(torch.autograd.Function):
    """J(q0, xy) of one solver; see flow_objective."""

    @staticmethod
    def forward(ctx, q0, xy, solver, n_steps, wx, wy, mem_cap_bytes):
        n = solver._n_cells * 4
        if q0.numel() != n:
            raise ValueError("q0 has %d values; the solver's state has %d" % (q0.numel(), n))
        q = q0.detach().to(device="cpu", dtype=torch.float64).contiguous().numpy().ravel()
        want_q = ctx.needs_input_grad[0]
        want_xy = xy is not None and ctx.needs_input_grad[1]
        if xy is not None:
            m = solver.mesh
            xh = xy.detach().to(device="cpu", dtype=torch.float64).numpy()
            if xh.shape != (m.n_nodes, 2) or not (np.array_equal(xh[:, 0], m.export("nodesX")) and
                                                    np.array_equal(xh[:, 1], m.export("nodesY"))):
                raise ValueError("xy must hold the solver mesh's node coordinates, shape (n_nodes, 2)")
        ctx.q_shape, ctx.q_dev, ctx.q_dtype = q0.shape, q0.device, q0.dtype
        ctx.xy_meta = (xy.shape, xy.device, xy.dtype) if xy is not None else None
        ctx.g_q = ctx.g_xy = None
        if want_q or want_xy:
            if want_xy:
                J, g, gxy, _ = solver.adjoint(q, n_steps, wx=wx, wy=wy, mem_cap_bytes=mem_cap_bytes, nodes=True)
                ctx.g_xy = gxy
            else:
                J, g, _ = solver.adjoint(q, n_steps, wx=wx, wy=wy, mem_cap_bytes=mem_cap_bytes)
            ctx.g_q = g
        else:
            J = solver.objective(q, n_steps, wx=wx, wy=wy)
        return torch.tensor(J, dtype=torch.float64, device=q0.device)

    @staticmethod
    def backward(ctx, gJ):
        s = gJ.detach().to("cpu", torch.float64)
        gq = gxy = None
        if ctx.g_q is not None and ctx.needs_input_grad[0]:
            gq = (torch.from_numpy(ctx.g_q).reshape(ctx.q_shape) * s).to(device=ctx.q_dev, dtype=ctx.q_dtype)
        if ctx.g_xy is not None and ctx.needs_input_grad[1]:
            shape, dev, dt = ctx.xy_meta
            gxy = (torch.from_numpy(ctx.g_xy).reshape(shape) * s).to(device=dev, dtype=dt)
        return gq, gxy, None, None, None, None, None


def flow_objective(solver, q0: torch.Tensor, n_steps: int, wx=None, wy=None, xy: torch.Tensor = None,
                   mem_cap_bytes: float = 8e9) -> torch.Tensor:
    """Objective J after n_steps steps from q0, differentiable in q0 (and in
    xy, the mesh's node coordinates, when given).  wx, wy: per-boundary-element
    weights of the pressure force, as Solver.objective takes them."""
    if solver.precision != "f64":
        raise ValueError("the adjoint is FP64 only")
    return FlowObjective.apply(q0, xy, solver, int(n_steps), wx, wy, float(mem_cap_bytes))
