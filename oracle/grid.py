"""Wavefield on a uniform grid (TEST INFRASTRUCTURE): a restatement of
``flat_shell_engine/spec2meshgrid_flat_shell.m:1-583``.

* Grid (``:103-109``): ``Nx`` x ``Ny`` points inset by 1e-8 m from the mesh's bounding box;
  point ``j = r + Ny*c`` lies at ``(xi[c], yi[r])`` (MATLAB ``meshgrid`` + column-major
  ``reshape``).  Coordinates are ``a + k*d``; MATLAB's colon operator can round the last
  bit of the last points differently, which can move a point on an element edge to the
  neighbouring owner: parity unpinned for such ties.
* Owner element (``:146-186``): among the searched elements (the leading ``n_search``;
  the reference leaves out the delaminated copies) the ones whose corner bounding box holds
  the point; if exactly one, it; else the first with all four edge cross products negative
  (strictly inside a counter-clockwise quad); else the first candidate.
* Local coordinates (``:188-299``): start at the owner's nearest element node (first index
  on ties) with that node's GLL coordinates, one step with the inverse Jacobian of the
  element's first node, then three Newton steps ``[xi, eta] += J^{-1} (x_p - x(xi, eta))``.
  The reference clamps xi, eta to [-1, 1] only after it formed its shape-function matrix,
  so the clamp has no effect, and none is applied here.
* Shape functions: the reference forms them from the inverse 2D Vandermonde matrix of
  monomials (``shape2Dp.m``, ``Vandermonde2D.m``); here the same polynomials as products of
  1D Lagrange polynomials on the GLL nodes (``lagrange_basis``; equality with the
  Vandermonde form is a test).
* Values (``:321-575``): the quantity ``motion`` 1-9 of ``field_variable`` at the top or
  bottom surface (``u + h/2 phi_x`` etc.); for the top surface the nodes under a
  delamination take the values of their twins above it (``:567-569``); then
  ``Z = sum_k N_k U[node_k]`` per point.  Result ``[n_frames, Ny, Nx]``, ``[n, r, c]`` at
  ``(xi[c], yi[r])``.
"""
from __future__ import annotations

import numpy as np

from .common import gll

EPS = 1e-8


def grid_points(coords, Nx, Ny):
    x, y = coords[:, 0], coords[:, 1]
    L = x.max() - x.min() - 2 * EPS
    B = y.max() - y.min() - 2 * EPS
    xi = x.min() + EPS + L / (Nx - 1) * np.arange(Nx)
    yi = y.min() + EPS + B / (Ny - 1) * np.arange(Ny)
    xp = np.repeat(xi, Ny)  # j = r + Ny*c
    yp = np.tile(yi, Nx)
    return xi, yi, xp, yp


def lagrange_basis(xn, t):
    """L[k, m] = L_m(t_k) for the Lagrange polynomials on nodes xn."""
    t = np.asarray(t, float).reshape(-1, 1)
    n = len(xn)
    L = np.ones((t.shape[0], n))
    for m in range(n):
        for q in range(n):
            if q != m:
                L[:, m] *= (t[:, 0] - xn[q]) / (xn[m] - xn[q])
    return L


def lagrange_basis_derivative(xn, t):
    t = np.asarray(t, float).reshape(-1)
    n = len(xn)
    dL = np.zeros((t.size, n))
    for m in range(n):
        for s in range(n):
            if s == m:
                continue
            term = np.full(t.size, 1.0 / (xn[m] - xn[s]))
            for q in range(n):
                if q != m and q != s:
                    term *= (t - xn[q]) / (xn[m] - xn[q])
            dL[:, m] += term
    return dL


def vandermonde_shape(n, ksi, x, y):
    """``shape2Dp.m`` with ``Vandermonde2D.m``: N[k, p] for local node k (xi fastest)."""
    V = np.zeros((n * n, n * n))
    r = 0
    for i1 in range(n):
        for i2 in range(n):
            c = 0
            for j1 in range(n):
                for j2 in range(n):
                    V[r, c] = ksi[i2] ** j2 * ksi[i1] ** j1
                    c += 1
            r += 1
    Q = np.linalg.inv(V)
    f = np.array([np.asarray(x) ** j2 * np.asarray(y) ** j1 for j1 in range(n) for j2 in range(n)])
    return Q.T @ f


def owner_elements(nodes, coords, n, n_search, xp, yp):
    mc = [0, n - 1, n * n - 1, n * n - n]  # counter-clockwise corners (:73)
    ne = np.arange(n_search)
    x = coords[nodes[ne][:, mc], 0].T  # [4, fen]
    y = coords[nodes[ne][:, mc], 1].T
    x21, x32, x43, x14 = x[1] - x[0], x[2] - x[1], x[3] - x[2], x[0] - x[3]
    y21, y32, y43, y14 = y[1] - y[0], y[2] - y[1], y[3] - y[2], y[0] - y[3]
    pmin_x, pmin_y, pmax_x, pmax_y = x.min(0), y.min(0), x.max(0), y.max(0)
    owner = np.zeros(xp.size, np.int64)
    for j in range(xp.size):
        I = np.nonzero((xp[j] >= pmin_x) & (xp[j] <= pmax_x) & (yp[j] >= pmin_y) & (yp[j] <= pmax_y))[0]
        if I.size == 1:
            owner[j] = I[0]
            continue
        d1 = (xp[j] * y21[I] - x[0, I] * y21[I]) - (yp[j] * x21[I] - y[0, I] * x21[I])
        d2 = (xp[j] * y32[I] - x[1, I] * y32[I]) - (yp[j] * x32[I] - y[1, I] * x32[I])
        d3 = (xp[j] * y43[I] - x[2, I] * y43[I]) - (yp[j] * x43[I] - y[2, I] * x43[I])
        d4 = (xp[j] * y14[I] - x[3, I] * y14[I]) - (yp[j] * x14[I] - y[3, I] * x14[I])
        inside = np.nonzero((d1 < 0) & (d2 < 0) & (d3 < 0) & (d4 < 0))[0]
        owner[j] = I[inside[0]] if inside.size else I[0]
    return owner


def local_coordinates(nodes, coords, n, owner, xp, yp, iterations=3):
    ksi, _ = gll(n)
    en = nodes[owner]                         # [P, n^2]
    xe, ye = coords[en, 0], coords[en, 1]
    k0 = np.argmin(np.hypot(xe - xp[:, None], ye - yp[:, None]), axis=1)
    rows = np.arange(owner.size)
    x0, y0 = xe[rows, k0], ye[rows, k0]
    s, t = ksi[k0 % n].copy(), ksi[k0 // n].copy()

    def jac(s, t):
        Ls, Lt = lagrange_basis(ksi, s), lagrange_basis(ksi, t)
        dLs, dLt = lagrange_basis_derivative(ksi, s), lagrange_basis_derivative(ksi, t)
        Nx = (dLs[:, None, :] * Lt[:, :, None]).reshape(-1, n * n)   # d/dxi, node k = j*n + i
        Ny = (Ls[:, None, :] * dLt[:, :, None]).reshape(-1, n * n)
        J11, J12 = (Nx * xe).sum(1), (Ny * xe).sum(1)
        J21, J22 = (Nx * ye).sum(1), (Ny * ye).sum(1)
        a = 1.0 / (J11 * J22 - J12 * J21)
        return a * J22, -a * J12, -a * J21, a * J11

    def position(s, t):
        N = shape(s, t)
        return (N * xe).sum(1), (N * ye).sum(1)

    def shape(s, t):
        Ls, Lt = lagrange_basis(ksi, s), lagrange_basis(ksi, t)
        return (Ls[:, None, :] * Lt[:, :, None]).reshape(-1, n * n)

    # first step with the inverse Jacobian at the element's first node (:214-217)
    i11, i12, i21, i22 = jac(np.full(owner.size, ksi[0]), np.full(owner.size, ksi[0]))
    s, t = s + i11 * (xp - x0) + i12 * (yp - y0), t + i21 * (xp - x0) + i22 * (yp - y0)
    x1, y1 = position(s, t)
    for _ in range(iterations):
        i11, i12, i21, i22 = jac(s, t)
        s, t = s + i11 * (xp - x1) + i12 * (yp - y1), t + i21 * (xp - x1) + i22 * (yp - y1)
        x1, y1 = position(s, t)
    return s, t, shape(s, t)


def surface_values(frame, motion, surface, h, nf=5):
    """Nodal quantity ``motion`` (1-9) of one frame [dof] (u, phix, v, phiy, w per node)."""
    U = np.asarray(frame, float).reshape(-1, nf)
    ux, fx, uy, fy, uz = U[:, 0], U[:, 1], U[:, 2], U[:, 3], U[:, 4]
    sg = 1.0 if surface == "top" else -1.0
    if motion <= 5:
        return [ux, uy, uz, fx, fy][motion - 1].copy()
    if motion == 6:
        return ux + sg * h / 2 * fx
    if motion == 7:
        return uy + sg * h / 2 * fy
    if motion == 8:
        return np.sqrt((ux + sg * h / 2 * fx) ** 2 + (uy + sg * h / 2 * fy) ** 2)
    return np.sqrt((ux + sg * h / 2 * fx) ** 2 + (uy + sg * h / 2 * fy) ** 2 + uz ** 2)


def spec2meshgrid(nodes, coords, n, frames, Nx, Ny, motion, surface="top", h=0.0, n_search=None,
                  den_under=(), den_above=()):
    nodes = np.asarray(nodes, np.int64)
    coords = np.asarray(coords, float)
    n_search = nodes.shape[0] if n_search is None else n_search
    xi, yi, xp, yp = grid_points(coords, Nx, Ny)
    owner = owner_elements(nodes, coords, n, n_search, xp, yp)
    _, _, N = local_coordinates(nodes, coords, n, owner, xp, yp)
    out = np.zeros((len(frames), Ny, Nx))
    for f, fr in enumerate(frames):
        U = surface_values(fr, motion, surface, h)
        if surface == "top" and len(den_above):
            U[nodes[np.asarray(den_under)].reshape(-1)] = U[nodes[np.asarray(den_above)].reshape(-1)]
        Z = (N * U[nodes[owner]]).sum(1)          # j = r + Ny*c
        out[f] = Z.reshape(Nx, Ny).T
    return dict(data=out, xi=xi, yi=yi, owner=owner)
