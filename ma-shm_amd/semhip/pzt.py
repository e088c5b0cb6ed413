"""PZT transducers on the host: constants, actuator equivalent forces and sensor voltage
functionals through the library's C-ABI (sem_pzt_*), in the shapes that
``Solver.set_excitation`` and ``Solver.set_outputs`` take.

Two ways to place a transducer on a 3D plate:

* meshed: the transducer's hexahedra are part of the mesh (the reference's solid engine,
  ``pztEl``); ``actuator(...)`` / ``sensor(...)`` take those elements directly;
* on the surface: ``surface_actuator`` / ``surface_sensor`` build the transducer's
  hexahedra above top-face elements of the plate only to form its electrical matrices,
  then carry every force (and every voltage weight) of a transducer node down its
  through-thickness column to the plate node under it (the transducer moves with the
  surface it is bonded to).  This keeps a thin transducer from setting the stable time
  step: a 0.2 mm PZT layer of order-4 hexahedra has a critical step near 8e-9 s, below
  the 2e-8 s of a 100 x 100 x 2 order-4 plate run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def gll_nodes(n):
    """(nodes, weights) of the n-point GLL rule (sem_gll)."""
    x = np.zeros(n); w = np.zeros(n)
    _check(L.load().sem_gll(int(n), L.ptr(x, C.c_double), L.ptr(w, C.c_double)), "sem_gll")
    return x, w


def _check(rc, what):
    if rc != 0:
        raise L.SemError(rc, what)


def pzt5a():
    """PZT-5A in stress-charge form (the common data-sheet set: c11 = 121 GPa, c12 = 75.4,
    c13 = 75.2, c33 = 111, c44 = 21.1, c66 = 22.6 GPa; e31 = -5.4, e33 = 15.8,
    e15 = 12.3 C/m^2; eps11 = 8.11e-9, eps33 = 7.35e-9 F/m; rho = 7750 kg/m^3).
    Returns (Q [6x6], e [3x6], g [3x3], rho) with strain order ex ey ez gxy gyz gxz, the row
    order of the coupling B matrix, and Q in the solid element's D layout."""
    c11, c12, c13, c33, c44, c66 = 121e9, 75.4e9, 75.2e9, 111e9, 21.1e9, 22.6e9
    Q = np.zeros((6, 6))
    Q[:3, :3] = [[c11, c12, c13], [c12, c11, c13], [c13, c13, c33]]
    Q[3, 3], Q[4, 4], Q[5, 5] = c66, c44, c44
    e = np.zeros((3, 6))
    e[2, :3] = [-5.4, -5.4, 15.8]
    e[0, 5] = 12.3  # D_x from gamma_xz
    e[1, 4] = 12.3  # D_y from gamma_yz
    g = np.diag([8.11e-9, 8.11e-9, 7.35e-9])
    return Q, e, g, 7750.0


def pzt5a_strain_charge():
    """PZT-5A strain-charge data sheet (columns ex ey ez gxz gyz gxy): s11 = 16.4, s12 = -5.74,
    s13 = -7.22, s33 = 18.8, s44 = 47.5, s66 = 2 (s11 - s12) [1e-12 m^2/N]; d31 = -171,
    d33 = 374, d15 = 584 [1e-12 C/N]; eps11 = 1730 eps0, eps33 = 1700 eps0; rho = 7750.
    Returns (S, d, epsT, rho)."""
    eps0 = 8.854187817e-12
    s11, s12, s13, s33, s44 = 16.4e-12, -5.74e-12, -7.22e-12, 18.8e-12, 47.5e-12
    S = np.array([[s11, s12, s13, 0, 0, 0], [s12, s11, s13, 0, 0, 0], [s13, s13, s33, 0, 0, 0],
                  [0, 0, 0, s44, 0, 0], [0, 0, 0, 0, s44, 0], [0, 0, 0, 0, 0, 2 * (s11 - s12)]])
    d = np.zeros((3, 6))
    d[0, 3] = d[1, 4] = 584e-12
    d[2, :3] = [-171e-12, -171e-12, 374e-12]
    return S, d, np.diag([1730 * eps0, 1730 * eps0, 1700 * eps0]), 7750.0


def solid_coef_row(Q, rho):
    """14-entry solid coefficient row (D11 D12 D13 D14 D22 D23 D24 D33 D34 D44 D55 D56 D66
    rho) of a transducer element (main_spec3Dcuda_global_..._mat.m:506-521)."""
    return np.array([Q[0, 0], Q[0, 1], Q[0, 2], Q[0, 3], Q[1, 1], Q[1, 2], Q[1, 3], Q[2, 2], Q[2, 3], Q[3, 3],
                     Q[4, 4], Q[4, 5], Q[5, 5], rho])


def pzt_const(S, d, epsT, variant="solid", theta_deg=0.0):
    """(Q, e, g, (e31, e32, g33)) from the strain-charge form (sem_pzt_const)."""
    lib = L.load()
    S = L.f64(S); d = L.f64(d); epsT = L.f64(epsT)
    Q = np.zeros((6, 6)); e = np.zeros((3, 6)); g = np.zeros((3, 3)); x = np.zeros(3)
    _check(lib.sem_pzt_const(1 if variant == "solid" else 0, L.ptr(S, C.c_double), L.ptr(d, C.c_double),
                             L.ptr(epsT, C.c_double), float(theta_deg), L.ptr(Q, C.c_double), L.ptr(e, C.c_double),
                             L.ptr(g, C.c_double), L.ptr(x, C.c_double)), "sem_pzt_const")
    return Q, e, g, tuple(x)


def actuator(pnodes, coords, n_xy, n_z, e, g, V0=1.0, frame=1):
    """Force vector [3 n_nodes] of a meshed transducer at voltage V0."""
    lib = L.load()
    pn = L.i32(pnodes); X = L.f64(coords); e = L.f64(e); g = L.f64(g)
    fa = np.zeros(3 * X.shape[0])
    _check(lib.sem_pzt_actuator(n_xy, n_z, pn.shape[0], L.ptr(pn, C.c_int32), X.shape[0], L.ptr(X, C.c_double),
                                L.ptr(e, C.c_double), L.ptr(g, C.c_double), int(frame), float(V0),
                                L.ptr(fa, C.c_double)), "sem_pzt_actuator")
    return fa


def sensor(pnodes, coords, n_xy, n_z, e, g, frame=1):
    """Voltage weights [3 n_nodes] of a meshed transducer (voltage = w . U)."""
    lib = L.load()
    pn = L.i32(pnodes); X = L.f64(coords); e = L.f64(e); g = L.f64(g)
    w = np.zeros(3 * X.shape[0])
    _check(lib.sem_pzt_sensor(n_xy, n_z, pn.shape[0], L.ptr(pn, C.c_int32), X.shape[0], L.ptr(X, C.c_double),
                              L.ptr(e, C.c_double), L.ptr(g, C.c_double), int(frame), L.ptr(w, C.c_double)),
           "sem_pzt_sensor")
    return w


def top_elements(nodes, coords, center, half, n_xy, n_z):
    """Plate elements whose top face lies at the top surface with its centre inside the
    square |x - cx| < half, |y - cy| < half."""
    nodes = np.asarray(nodes)
    nf = n_xy * n_xy
    ztop = coords[:, 2].max()
    face = nodes[:, (n_z - 1) * nf:]
    zf = coords[face, 2]
    on_top = np.all(np.isclose(zf, ztop), axis=1)
    c = coords[face].mean(axis=1)
    inside = (np.abs(c[:, 0] - center[0]) < half) & (np.abs(c[:, 1] - center[1]) < half)
    return np.nonzero(on_top & inside)[0]


def surface_transducer(nodes, coords, n_xy, n_z, elems, t_pzt):
    """Transducer hexahedra above the top faces of plate elements ``elems``: bottom level =
    the plate's top-face nodes, level k at z_top + (xi_k + 1)/2 t_pzt.  Returns
    (pnodes, coords_all, column) where column[v] is the plate node under node v."""
    nodes = np.asarray(nodes)
    nf = n_xy * n_xy
    face = nodes[elems][:, (n_z - 1) * nf:]
    bot_u, J = np.unique(face, return_inverse=True)
    J = J.reshape(face.shape)
    N0 = coords.shape[0]
    xz = gll_nodes(n_z)[0]
    pn = np.zeros((len(elems), nf * n_z), np.int64)
    pn[:, :nf] = face
    col = np.arange(N0 + (n_z - 1) * bot_u.size)
    X = np.zeros((col.size, 3))
    X[:N0] = coords
    for k in range(1, n_z):
        ids = N0 + (k - 1) * bot_u.size + np.arange(bot_u.size)
        col[ids] = bot_u
        X[ids] = coords[bot_u]
        X[ids, 2] += (xz[k] + 1) / 2 * t_pzt
        pn[:, k * nf:(k + 1) * nf] = N0 + (k - 1) * bot_u.size + J
    return pn, X, col


def _condense(v, col, n_nodes):
    out = np.zeros(3 * n_nodes)
    v = v.reshape(-1, 3)
    np.add.at(out.reshape(-1, 3), col, v)
    return out


def surface_actuator(nodes, coords, n_xy, n_z, elems, t_pzt, e, g, V0=1.0):
    """(dofs, amp) of a surface transducer's forces at voltage V0 on the plate nodes."""
    pn, X, col = surface_transducer(nodes, coords, n_xy, n_z, elems, t_pzt)
    f = _condense(actuator(pn, X, n_xy, n_z, e, g, V0, frame=1), col, coords.shape[0])
    dofs = np.nonzero(f)[0]
    return dofs, f[dofs]


def surface_sensor(nodes, coords, n_xy, n_z, elems, t_pzt, e, g):
    """(dofs, weights) of a surface transducer's open-circuit voltage functional."""
    pn, X, col = surface_transducer(nodes, coords, n_xy, n_z, elems, t_pzt)
    w = _condense(sensor(pn, X, n_xy, n_z, e, g, frame=1), col, coords.shape[0])
    dofs = np.nonzero(w)[0]
    return dofs, w[dofs]


def shell_actuator(act_nodes, coords, n_xy, n_z, t_pzt, h, e, g, V0=1.0):
    """(dofs, amp) of a flat-shell actuator (forces and moments on its footprint boundary)."""
    lib = L.load()
    an = L.i32(act_nodes); X = L.f64(coords); e = L.f64(e); g = L.f64(g)
    fa = np.zeros(5 * X.shape[0])
    _check(lib.sem_pzt_shell_actuator(n_xy, n_z, an.shape[0], L.ptr(an, C.c_int32), X.shape[0],
                                      L.ptr(X, C.c_double), float(t_pzt), float(h), L.ptr(e, C.c_double),
                                      L.ptr(g, C.c_double), float(V0), L.ptr(fa, C.c_double)),
           "sem_pzt_shell_actuator")
    dofs = np.nonzero(fa)[0]
    return dofs, fa[dofs]


def laminate_shell_add_pzt(row, h, t_pzt, rho_pzt, Q):
    lib = L.load()
    r = L.f64(row); Q = L.f64(Q); out = np.zeros(23)
    _check(lib.sem_laminate_shell_add_pzt(L.ptr(r, C.c_double), float(h), float(t_pzt), float(rho_pzt),
                                          L.ptr(Q, C.c_double), L.ptr(out, C.c_double)), "sem_laminate_shell_add_pzt")
    return out


def shell_sensor(solver, elems, t_pzt, h, e31, e32, g33):
    """(dofs, weights) of a flat-shell sensor's voltage over the uploaded mesh."""
    el = L.i32(elems)
    w = np.zeros(solver.dof)
    solver._c(solver._lib.sem_shell_pzt_sensor(solver._h, el.size, L.ptr(el, C.c_int32), float(t_pzt), float(h),
                                               float(e31), float(e32), float(g33), L.ptr(w, C.c_double)))
    dofs = np.nonzero(w)[0]
    return dofs, w[dofs]
