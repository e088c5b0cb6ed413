"""Discrete adjoint of the central-difference loop (TEST INFRASTRUCTURE).

The reference integrates forward only; its loop (``force_excit_c.m:686-693``, solid engine
``main_spec3Dcuda_global_cpu_gpu3_wave_sens_opt_single_PCT_mat.m:1032-1036``) is, per 1-based step n,

    R_n = M (u_{n+1} - 2 u_n + u_{n-1}) + dt/2 C (u_{n+1} - u_{n-1}) + dt^2 (K u_n - f st(n)) = 0,

C = dv .* M.  For a functional J of the traces, with K = sum_e s_e K_e0 and
M = sum_e r_e M_e0 (per-element stiffness / density scales), the Lagrangian
J - sum_n mu_n^T R_n gives, for n = nn_last .. nn_first with mu_{nn_last+1} = mu_{nn_last+2} = 0,

    (M + dt/2 C) mu_n = dJ/du_{n+1} + (2 M - dt^2 K) mu_{n+1} - (M - dt/2 C) mu_{n+2}
    dJ/ds_e = -dt^2 sum_n mu_n^T K_e0 u_n
    dJ/dr_e = -sum_n mu_n^T M_e0 [(u_{n+1} - 2 u_n + u_{n-1}) + dt/2 dv (u_{n+1} - u_{n-1})].

This module computes those in NumPy/FP64 on the oracle's ShellModel / SolidModel.  It is
pinned by central finite differences of the oracle's own forward loop
(``tests/test_oracle_adjoint.py``), for every trace quantity, with damping, on shells and
solids.

Traces: output j at step nn is sum_t w_t q(u)[dof_t], q = u_nn ('disp'), u_{nn+1}
('disp_post'), (u_{nn+1} - u_{nn-1}) / (2 dt) ('vel') or (u_{nn+1} - 2 u_nn + u_{nn-1}) / dt^2
('acc'), as sampled by the product's kernels (force_excit_c.m:701-705 for the point case).
"""
from __future__ import annotations

import numpy as np

QUANTITIES = ("disp", "vel", "disp_post", "acc")


def forward_states(model, dt, st, f_ext, nn_first, nn_last, u0=None, uold0=None):
    """States u_{nn_first-1} .. u_{nn_last+1} ([nsteps+2 x dof]) of the loop
    ``force_excit_c.m:686-693``: F = -K U + 2 a0 M U + f st(nn) - a0 M uold + a1 C uold,
    unew = F / (a0 M + a1 C)."""
    a0 = 1.0 / dt**2
    a1 = 1.0 / (2 * dt)
    mg = model.mass
    cg = model.dv * mg
    mg0 = 1.0 / (a0 * mg + a1 * cg)
    U = np.zeros(model.dof) if u0 is None else np.array(u0, dtype=np.float64)
    uold = np.zeros(model.dof) if uold0 is None else np.array(uold0, dtype=np.float64)
    out = [uold, U]
    for nn in range(nn_first, nn_last + 1):
        F = -model.internal_force(U) + 2 * a0 * mg * U + f_ext * st[nn - 1] - a0 * mg * uold + a1 * cg * uold
        unew = mg0 * F
        out.append(unew)
        uold, U = U, unew
    return np.array(out)


def traces_from_states(states, nn_first, nn_last, ptr, dofs, w, quantity, dt):
    """Traces [nn_last x n_out] (row nn-1) of functionals ptr/dofs/w from forward_states."""
    n_out = len(ptr) - 1
    tr = np.zeros((nn_last, n_out))
    a0 = 1.0 / dt**2
    a1 = 1.0 / (2 * dt)
    for nn in range(nn_first, nn_last + 1):
        m = nn - nn_first + 1  # states[m] = u_nn
        uo, u, un = states[m - 1], states[m], states[m + 1]
        if quantity == "disp":
            q = u
        elif quantity == "disp_post":
            q = un
        elif quantity == "vel":
            q = a1 * (un - uo)
        else:
            q = a0 * (uo - 2 * u + un)
        for j in range(n_out):
            sl = slice(ptr[j], ptr[j + 1])
            tr[nn - 1, j] = np.dot(w[sl], q[dofs[sl]])
    return tr


def _point_mass(model_unit):
    """Element-local lumped mass at every point [L x nf] of a model built with unit scales
    (``force_excit_c.m:497-503``; ``mass3D_v2.m:51-57``)."""
    if model_unit.nf == 5:
        W = model_unit.WWDetJ
        return np.stack([W * model_unit.m11, W * model_unit.I11, W * model_unit.m11, W * model_unit.I11,
                         W * model_unit.m11], axis=1)
    W = model_unit.WWWDetJ * model_unit.rho
    return np.stack([W, W, W], axis=1)


def adjoint_gradients(model, model_unit, dt, st, f_ext, nn_first, nn_last, ptr, dofs, w, quantity, dJ,
                      u0=None, uold0=None):
    """(dJ/ds [fen], dJ/dr [fen]) for J with dJ/dtraces = dJ [>= nn_last x n_out].

    ``model`` carries the scales of the run, ``model_unit`` the same mesh and material with
    unit scales (its local forces are K_e0 u, its point masses M_e0)."""
    nf = model.nf
    a0 = 1.0 / dt**2
    a1 = 1.0 / (2 * dt)
    states = forward_states(model, dt, st, f_ext, nn_first, nn_last, u0, uold0)
    ptr = np.asarray(ptr, np.int64)
    dofs = np.asarray(dofs, np.int64)
    w = np.asarray(w, np.float64)
    dJ = np.asarray(dJ, np.float64).reshape(dJ.shape[0], -1)
    n_out = len(ptr) - 1

    def d(nn):
        return dJ[nn - 1] if nn_first <= nn <= nn_last else np.zeros(n_out)

    def g(k):
        """dJ/du_k."""
        if quantity == "disp":
            s = d(k)
        elif quantity == "disp_post":
            s = d(k - 1)
        elif quantity == "vel":
            s = a1 * (d(k - 1) - d(k + 1))
        else:
            s = a0 * (d(k - 1) - 2 * d(k) + d(k + 1))
        v = np.zeros(model.dof)
        for j in range(n_out):
            sl = slice(ptr[j], ptr[j + 1])
            np.add.at(v, dofs[sl], w[sl] * s[j])
        return v

    M = model.mass
    Cd = model.dv * M
    A = M + 0.5 * dt * Cd
    B = M - 0.5 * dt * Cd
    I = model.nodes.reshape(-1)
    pidx = (nf * I)[:, None] + np.arange(nf)[None, :]  # [L x nf] DOF of every local point
    mq = _point_mass(model_unit)
    dvq = np.asarray(model.dv[:nf])
    gs = np.zeros(model.fen)
    gr = np.zeros(model.fen)
    mu1 = np.zeros(model.dof)  # mu_{n+1}
    mu2 = np.zeros(model.dof)  # mu_{n+2}
    for n in range(nn_last, nn_first - 1, -1):
        mu = (g(n + 1) + 2 * M * mu1 - dt**2 * model.internal_force(mu1) - B * mu2) / A
        m = n - nn_first + 1
        uo, u, un = states[m - 1], states[m], states[m + 1]
        lam = mu[pidx]
        ke = model_unit.local_forces(u)
        gs -= dt**2 * (lam * ke).reshape(model.fen, -1).sum(axis=1)
        acc = (un - 2 * u + uo)[pidx] + 0.5 * dt * dvq[None, :] * (un - uo)[pidx]
        gr -= (lam * mq * acc).reshape(model.fen, -1).sum(axis=1)
        mu2, mu1 = mu1, mu
    return gs, gr
