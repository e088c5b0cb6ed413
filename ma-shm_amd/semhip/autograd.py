"""torch.autograd binding of the time loop: traces as a differentiable function of the
per-element stiffness and density scales.

The forward pass runs the HIP time loop (``Solver.run``); the backward pass runs the
library's discrete adjoint (``Solver.adjoint``: checkpointed forward sweep, reverse
central-difference sweep through the same element kernels, per-element gradient
epilogue).  Both start from the same initial state, set before each pass.  Tensors
cross the C-ABI as host float64 arrays; gradients come back on the device and dtype of
the scale tensors.
"""
from __future__ import annotations

import numpy as np
import torch


class _TraceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, stiff_scale, rho_scale, sim):
        sim._prepare(stiff_scale, rho_scale)
        r = sim.solver.run(sim.dt, sim.nn_first, sim.nn_last, sim.precision, want_frames=False)
        if r["status"] != "ok":
            raise RuntimeError(f"time loop diverged at step {r['nn_done']}")
        ctx.sim = sim
        ctx.ss = stiff_scale.detach().cpu().numpy().astype(np.float64)
        ctx.rs = rho_scale.detach().cpu().numpy().astype(np.float64)
        ctx.meta = (stiff_scale.device, stiff_scale.dtype)
        tr = r["traces"][sim.nn_first - 1:sim.nn_last]
        return torch.as_tensor(np.ascontiguousarray(tr), device=stiff_scale.device, dtype=stiff_scale.dtype)

    @staticmethod
    def backward(ctx, grad):
        sim = ctx.sim
        device, dtype = ctx.meta
        sim._prepare_np(ctx.ss, ctx.rs)
        dJ = np.zeros((sim.nn_last, sim.solver.n_out))
        dJ[sim.nn_first - 1:] = grad.detach().cpu().numpy().astype(np.float64)
        g = sim.solver.adjoint(sim.dt, sim.nn_first, sim.nn_last, dJ, sim.precision, sim.seg_len)
        gs = torch.as_tensor(g["g_stiff"], device=device, dtype=dtype)
        gr = torch.as_tensor(g["g_rho"], device=device, dtype=dtype)
        return gs, gr, None


class TraceSimulation:
    """A configured ``Solver`` (mesh, excitation, outputs, damping already set) plus the
    material table and the step range.  ``traces(stiff_scale, rho_scale)`` returns the
    output traces [nn_last - nn_first + 1, n_out] and is differentiable in both scales."""

    def __init__(self, solver, coef, dt, nn_first, nn_last, elem_mat=None, u0=None, uold0=None,
                 precision="fp64", seg_len=0):
        self.solver = solver
        self.coef = np.asarray(coef, np.float64)
        self.elem_mat = elem_mat
        self.dt = float(dt)
        self.nn_first = int(nn_first)
        self.nn_last = int(nn_last)
        self.u0 = u0
        self.uold0 = uold0
        self.precision = precision
        self.seg_len = int(seg_len)

    def _prepare_np(self, ss, rs):
        self.solver.set_material(self.coef, self.elem_mat, ss, rs)
        self.solver.set_state(self.u0, self.uold0)

    def _prepare(self, stiff_scale, rho_scale):
        self._prepare_np(stiff_scale.detach().cpu().numpy().astype(np.float64),
                         rho_scale.detach().cpu().numpy().astype(np.float64))

    def traces(self, stiff_scale: torch.Tensor, rho_scale: torch.Tensor) -> torch.Tensor:
        return _TraceFn.apply(stiff_scale, rho_scale, self)
