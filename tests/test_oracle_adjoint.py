"""Pins of the oracle's discrete adjoint (oracle/adjoint.py) on the CPU.

The adjoint gradients must equal central finite differences of the oracle's forward loop
for every trace quantity, with damping, for stiffness and density scales, on a shell and
a solid mesh; the loop that the adjoint differentiates must reproduce the oracle's run().
"""
import numpy as np
import pytest

from cases import shell_pair, solid_pair
from oracle import adjoint as oa
from oracle import shell as osh
from oracle import solid as oso


def _functionals(dof, nf, rng):
    """One point output and one three-term functional."""
    d0 = nf * (dof // nf // 2) + nf - 1
    ds = rng.choice(dof, 3, replace=False)
    ptr = np.array([0, 1, 4])
    dofs = np.concatenate([[d0], ds])
    w = np.concatenate([[1.0], rng.uniform(0.5, 1.5, 3)])
    return ptr, dofs, w


def _J(model, dt, st, f, n0, n1, ptr, dofs, w, q, dJ, u0, uold0):
    S = oa.forward_states(model, dt, st, f, n0, n1, u0, uold0)
    return float((oa.traces_from_states(S, n0, n1, ptr, dofs, w, q, dt) * dJ[:n1]).sum())


def _check(make, nf, dt, nsteps, quantities, elems, seed):
    rng = np.random.default_rng(seed)
    model = make(None, None)
    unit = make(np.ones(model.fen), np.ones(model.fen))
    ss = rng.uniform(0.8, 1.2, model.fen)
    rs = rng.uniform(0.8, 1.2, model.fen)
    model = make(ss, rs)
    st = rng.standard_normal(nsteps + 5)
    f = np.zeros(model.dof)
    f[rng.choice(model.dof, 2, replace=False)] = rng.uniform(0.5, 1.0, 2)
    ptr, dofs, w = _functionals(model.dof, nf, rng)
    n0, n1 = 3, nsteps
    u0 = 1e-9 * rng.standard_normal(model.dof)
    uold0 = 1e-9 * rng.standard_normal(model.dof)
    for q in quantities:
        dJ = rng.standard_normal((n1, len(ptr) - 1))
        gs, gr = oa.adjoint_gradients(model, unit, dt, st, f, n0, n1, ptr, dofs, w, q, dJ, u0, uold0)
        h = 1e-4
        for e in elems:
            for which, g in (("s", gs), ("r", gr)):
                vals = []
                for sgn in (1, -1):
                    s2, r2 = ss.copy(), rs.copy()
                    (s2 if which == "s" else r2)[e] += sgn * h
                    vals.append(_J(make(s2, r2), dt, st, f, n0, n1, ptr, dofs, w, q, dJ, u0, uold0))
                fd = (vals[0] - vals[1]) / (2 * h)
                scale = max(np.abs(gs).max() if which == "s" else np.abs(gr).max(), 1e-300)
                assert abs(fd - g[e]) <= 1e-6 * scale, (q, which, e, fd, g[e])
    return gs, gr


def test_forward_states_reproduce_run():
    model, _ = shell_pair(nex=3, ney=2, order=3, dv=(2e3, 0, 2e3, 0, 1e3))
    rng = np.random.default_rng(1)
    st = rng.standard_normal(40)
    f = np.zeros(model.dof)
    f[5 * 7 + 4] = 1.0
    outs = [5 * 3 + 4, 5 * 11]
    ptr, dofs, w = np.array([0, 1, 2]), np.array(outs), np.ones(2)
    for q in ("disp", "vel", "disp_post"):
        o = model.run(1e-7, st, f, 1, 40, outs, q)
        S = oa.forward_states(model, 1e-7, st, f, 1, 40)
        tr = oa.traces_from_states(S, 1, 40, ptr, dofs, w, q, 1e-7)
        np.testing.assert_array_equal(tr, o["traces"][:40])
        np.testing.assert_array_equal(S[-1], o["U"])


def test_shell_adjoint_matches_finite_differences():
    def make(ss, rs):
        return shell_pair(nex=3, ney=2, order=3, dv=(3e3, 0, 3e3, 0, 2e3), stiff_scale=ss, rho_scale=rs)[0]
    _check(make, 5, 1e-7, 30, oa.QUANTITIES, (0, 4), seed=2)


def test_solid_adjoint_matches_finite_differences():
    def make(ss, rs):
        return solid_pair(nex=2, ney=1, nlay=1, order=2, order_z=2, dv=(2e3, 2e3, 1e3),
                          stiff_scale=ss, rho_scale=rs)[0]
    _check(make, 3, 2e-8, 25, ("disp_post", "acc"), (0, 1), seed=3)


def test_adjoint_is_linear_in_dJ():
    """Gradients are linear in dJ/dtraces: a plausible slip (a dropped source term or a
    wrong step index) keeps FD agreement for one quantity but not this identity across
    two independent dJ."""
    model, _ = shell_pair(nex=2, ney=2, order=3, dv=(1e3, 0, 1e3, 0, 1e3))
    unit = model
    rng = np.random.default_rng(5)
    st = rng.standard_normal(30)
    f = np.zeros(model.dof)
    f[4] = 1.0
    ptr, dofs, w = _functionals(model.dof, 5, rng)
    d1 = rng.standard_normal((20, 2))
    d2 = rng.standard_normal((20, 2))
    g1 = oa.adjoint_gradients(model, unit, 1e-7, st, f, 1, 20, ptr, dofs, w, "vel", d1)
    g2 = oa.adjoint_gradients(model, unit, 1e-7, st, f, 1, 20, ptr, dofs, w, "vel", d2)
    g12 = oa.adjoint_gradients(model, unit, 1e-7, st, f, 1, 20, ptr, dofs, w, "vel", 2 * d1 - d2)
    for k in range(2):
        np.testing.assert_allclose(g12[k], 2 * g1[k] - g2[k], rtol=1e-9, atol=1e-12 * np.abs(g1[k]).max())
