"""PZT transducers (TEST INFRASTRUCTURE): NumPy restatement of the reference's set-up.

* ``pzt_const_solid``  -- ``solid3D/pzt_const.m:1-25``
* ``pzt_const_shell``  -- ``flat_shell/flat_shell_engine/pzt_const.m:1-37``
* ``coupl3d``          -- ``solid3D/coupl3D_v2.m`` (frame=1) / ``flat_shell_engine/coupl3D_global.m``
                          (frame=0), evaluated with ``shape3D_prim_v2.m`` (monomials x inv(V),
                          1e-12 zeroing) exactly as coded
* ``electrical``       -- assembly of Kfiu / Kfifi and the electrode sets
                          (``main_spec3Dcuda_global_cpu_gpu3_wave_sens_opt_single_PCT_mat.m:198-255``)
* ``actuator_forces``  -- induced potential and FaL = Kfiu^T FiA (same file :244-255)
* ``sensor_voltage``   -- open-circuit voltage of a displacement field (same file :1056-1067)
* ``connect_pzt3d_to_plate``, ``find_boundary_nodes``, ``shell_actuator`` --
                          ``main_flat_shell_multi_pzt_c.m:158-235`` with its mesher helpers
* ``shell_sensor_voltage`` -- ``main_flat_shell_multi_pzt_c.m:618-623``
* ``laminate_shell_add_pzt`` -- ``composite_plate_c_delam_pzt.m:175-215``

Pins (tests/test_pzt.py): the strain-charge -> stress-charge conversion reproduces the
published PZT-5A stress-charge constants from its published strain-charge ones; on an
axis-aligned block under V0 the induced potential is linear through the thickness and the
forces are self-equilibrated with edge resultants e31 V0 L; under a uniform in-plane
strain the open-circuit voltage is e31 eps t / g33; the shell's footprint boundary is its
perimeter; with shear constants that all differ, one brick under each pure shear strain and
a linear potential gives u^T kufi phi = V gamma ep[i, c] with the column c that
``pzt_const.m:20-22`` pairs with ``coupl3D_v2.m``'s strain row (a swapped shear pair --
planted in a check of this test -- fails).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from .common import gll, vandermonde

PERM = [0, 1, 2, 5, 4, 3]  # (ex ey ez gxz gyz gxy) -> (ex ey ez gxy gxz gyz)


def pzt_const_solid(S, d, epsT):
    S = np.asarray(S, float); d = np.asarray(d, float); epsT = np.asarray(epsT, float)
    ep = d @ np.linalg.inv(S)
    g = epsT - ep @ d.T
    Q = np.linalg.inv(S[np.ix_(PERM, PERM)])
    e = ep[:, PERM]
    return Q, e, g


def pzt_const_shell(S, d, epsT, theta_deg):
    S = np.asarray(S, float); d = np.asarray(d, float); epsT = np.asarray(epsT, float)
    th = np.deg2rad(theta_deg)
    m, n = np.cos(th), np.sin(th)
    Tinv = np.array([[m * m, n * n, 0, 0, 0, -2 * m * n],
                     [n * n, m * m, 0, 0, 0, 2 * m * n],
                     [0, 0, 1, 0, 0, 0],
                     [0, 0, 0, m, n, 0],
                     [0, 0, 0, -n, m, 0],
                     [m * n, -m * n, 0, 0, 0, m * m - n * n]])
    dp_ = Tinv @ d.T
    ep = dp_.T @ np.linalg.inv(S)
    Q = np.linalg.inv(S)
    e = ep[:, PERM]
    return Q, e, epsT.copy(), ep[2, 0], ep[2, 1], epsT[2, 2]


def _shape3d_prim_v2(nx, ny, nz, Qx, Qy, Qz, x, y, z):
    """Derivatives of all shape functions at one point (``shape3D_prim_v2.m``)."""
    ux = x ** np.arange(nx); uy = y ** np.arange(ny); uz = z ** np.arange(nz)
    Sx, Sy, Sz = Qx @ ux, Qy @ uy, Qz @ uz
    dux = np.array([j * x ** (j - 1) if j else 0.0 for j in range(nx)])
    duy = np.array([j * y ** (j - 1) if j else 0.0 for j in range(ny)])
    duz = np.array([j * z ** (j - 1) if j else 0.0 for j in range(nz)])
    Sxp, Syp, Szp = Qx @ dux, Qy @ duy, Qz @ duz
    Nx = np.einsum("k,j,i->kji", Sz, Sy, Sxp).reshape(-1)
    Ny = np.einsum("k,j,i->kji", Sz, Syp, Sx).reshape(-1)
    Nz = np.einsum("k,j,i->kji", Szp, Sy, Sx).reshape(-1)
    for a in (Nx, Ny, Nz):
        a[np.abs(a) < 1e-12] = 0.0
    return Nx, Ny, Nz


def coupl3d(e, g, X, nx, nz, frame=1):
    """(kufi [3np x np], kfifi [np x np]) of one element; X [np x 3], local order x fastest."""
    ksi, wx = gll(nx)
    dz, wz = gll(nz)
    Qx = vandermonde(ksi)
    Qz = vandermonde(dz)
    np_ = nx * nx * nz
    kufi = np.zeros((3 * np_, np_))
    kfifi = np.zeros((np_, np_))
    for kz in range(nz):
        for ky in range(nx):
            for kx in range(nx):
                Nx, Ny, Nz = _shape3d_prim_v2(nx, nx, nz, Qx, Qx, Qz, ksi[kx], ksi[ky], dz[kz])
                J = np.array([Nx @ X, Ny @ X, Nz @ X])  # J[r, c] = d x_c / d xi_r
                v3 = np.cross(J[0], J[1]); v3 /= np.linalg.norm(v3)
                v1 = J[0] / np.linalg.norm(J[0])
                v2 = np.cross(v3, v1); v2 /= np.linalg.norm(v2)
                inJ = np.linalg.inv(J)
                w = wx[kx] * wx[ky] * wz[kz] * np.linalg.det(J)
                Ng = np.stack([Nx * inJ[0, 0] + Ny * inJ[0, 1] + Nz * inJ[0, 2],
                               Nx * inJ[1, 0] + Ny * inJ[1, 1] + Nz * inJ[1, 2],
                               Nx * inJ[2, 0] + Ny * inJ[2, 1] + Nz * inJ[2, 2]])  # [3 x np]
                B = np.zeros((6, 3 * np_))
                Bfi = np.zeros((3, np_))
                for cc in range(np_):
                    nxv, nyv, nzv = Ng[:, cc]
                    v11, v12, v13 = v1; v21, v22, v23 = v2; v31, v32, v33 = v3
                    if frame:
                        rows = [
                            [nxv * v11 ** 2 + nyv * v11 * v12 + nzv * v11 * v13,
                             nxv * v11 * v12 + nyv * v12 ** 2 + nzv * v12 * v13,
                             nxv * v11 * v13 + nyv * v12 * v13 + nzv * v13 ** 2],
                            [nxv * v21 ** 2 + nyv * v21 * v22 + nzv * v21 * v23,
                             nxv * v21 * v22 + nyv * v22 ** 2 + nzv * v22 * v23,
                             nxv * v21 * v23 + nyv * v22 * v23 + nzv * v23 ** 2],
                            [nxv * v31 ** 2 + nyv * v31 * v32 + nzv * v31 * v33,
                             nxv * v31 * v32 + nyv * v32 ** 2 + nzv * v32 * v33,
                             nxv * v31 * v33 + nyv * v32 * v33 + nzv * v33 ** 2],
                            [2 * nxv * v11 * v21 + nyv * (v12 * v21 + v11 * v22) + nzv * (v13 * v21 + v11 * v23),
                             2 * nyv * v12 * v22 + nxv * (v12 * v21 + v11 * v22) + nzv * (v13 * v22 + v12 * v23),
                             2 * nzv * v13 * v23 + nxv * (v13 * v21 + v11 * v23) + nyv * (v13 * v22 + v12 * v23)],
                            [2 * nxv * v21 * v31 + nyv * (v22 * v31 + v21 * v32) + nzv * (v23 * v31 + v21 * v33),
                             2 * nyv * v22 * v32 + nxv * (v22 * v31 + v21 * v32) + nzv * (v23 * v32 + v22 * v33),
                             2 * nzv * v23 * v33 + nxv * (v23 * v31 + v21 * v33) + nyv * (v23 * v32 + v22 * v33)],
                            [2 * nxv * v11 * v31 + nyv * (v12 * v31 + v11 * v32) + nzv * (v13 * v31 + v11 * v33),
                             2 * nyv * v12 * v32 + nxv * (v12 * v31 + v11 * v32) + nzv * (v13 * v32 + v12 * v33),
                             2 * nzv * v13 * v33 + nxv * (v13 * v31 + v11 * v33) + nyv * (v13 * v32 + v12 * v33)]]
                    else:
                        rows = [[nxv, 0, 0], [0, nyv, 0], [0, 0, nzv], [nyv, nxv, 0], [0, nzv, nyv], [nzv, 0, nxv]]
                    B[:, 3 * cc:3 * cc + 3] = np.array(rows)
                    Bfi[:, cc] = [nxv * v11 + nyv * v12 + nzv * v13, nxv * v21 + nyv * v22 + nzv * v23,
                                  nxv * v31 + nyv * v32 + nzv * v33]
                kufi += (B.T @ np.asarray(e).T @ Bfi) * w
                kfifi += (Bfi.T @ np.asarray(g) @ Bfi) * w
    return kufi, kfifi


def electrical(pnodes, coords, e, g, nx, nz, frame=1):
    """Assembled (Kfiu [n x 3n], Kfifi [n x n], EBC0, EBCI, EBCN, ids) over the transducer's
    own n nodes (ids = its node ids in ascending order; electrode sets as compact indices)."""
    pnodes = np.asarray(pnodes, np.int64)
    ids, loc = np.unique(pnodes, return_inverse=True)
    loc = loc.reshape(pnodes.shape)
    n = ids.size
    nf = nx * nx
    Kfiu = np.zeros((n, 3 * n))
    Kfifi = np.zeros((n, n))
    for el, lc in zip(pnodes, loc):
        kufi, kfifi = coupl3d(e, g, coords[el], nx, nz, frame)
        dofs = (3 * lc[:, None] + np.arange(3)[None, :]).reshape(-1)
        Kfiu[np.ix_(lc, dofs)] += kufi.T
        Kfifi[np.ix_(lc, lc)] += kfifi
    ebc0 = np.unique(loc[:, :nf])
    ebci = np.unique(loc[:, nf:-nf])
    ebcn = np.unique(loc[:, -nf:])
    return Kfiu, Kfifi, ebc0, ebci, ebcn, ids


def _dofs(ids):
    return (3 * np.asarray(ids)[:, None] + np.arange(3)[None, :]).reshape(-1)


def actuator_forces(pnodes, coords, e, g, nx, nz, V0, frame=1, return_potential=False):
    """Fa [3 N] over all N nodes (zero off the transducer)."""
    Kfiu, Kfifi, e0, ei, en, ids = electrical(pnodes, coords, e, g, nx, nz, frame)
    fiAi = -sla.solve(Kfifi[np.ix_(ei, ei)], Kfifi[np.ix_(ei, en)] @ np.full(en.size, V0))
    FiA = np.zeros(ids.size)
    FiA[en] = V0
    FiA[ei] = fiAi
    fa = np.zeros(3 * coords.shape[0])
    fa[_dofs(ids)] = Kfiu.T @ FiA
    if return_potential:
        phi = np.zeros(coords.shape[0])
        phi[ids] = FiA
        return fa, phi
    return fa


def sensor_voltage(pnodes, coords, e, g, nx, nz, U, frame=1):
    """Open-circuit voltage of the displacement field U [3 N]."""
    Kfiu, Kfifi, e0, ei, en, ids = electrical(pnodes, coords, e, g, nx, nz, frame)
    O = np.concatenate([ei, en])
    BN = np.concatenate([_dofs(s) for s in (e0, ei, en)])
    Ut = np.asarray(U)[_dofs(ids)]
    fio = -sla.solve(Kfifi[np.ix_(O, O)], Kfiu[np.ix_(O, BN)] @ Ut[BN])
    return float(np.mean(fio[-nx * nx:]))


def sensor_weights(pnodes, coords, e, g, nx, nz, frame=1):
    """w [3 N] with voltage = w . U: the voltage of every unit displacement of a transducer
    DOF, from the same open-circuit solve (all right-hand sides at once)."""
    Kfiu, Kfifi, e0, ei, en, ids = electrical(pnodes, coords, e, g, nx, nz, frame)
    O = np.concatenate([ei, en])
    BN = np.concatenate([_dofs(s) for s in (e0, ei, en)])
    fio = -sla.solve(Kfifi[np.ix_(O, O)], Kfiu[np.ix_(O, BN)])
    wt = np.zeros(3 * ids.size)
    wt[BN] = fio[-nx * nx:].mean(axis=0)
    w = np.zeros(3 * coords.shape[0])
    w[_dofs(ids)] = wt
    return w


def connect_pzt3d_to_plate(nx, nz, act_nodes, coords, t_pzt, h):
    """Virtual transducer elements over shell elements (``connect_pzt3D_to_plate_fun.m``):
    level 0 keeps the shell node ids, higher levels get ids after max(node) in the order
    of the sorted unique bottom nodes; returns (pnodes, coords_virtual)."""
    act_nodes = np.asarray(act_nodes, np.int64)
    nf = nx * nx
    ksi, _ = gll(nz)
    bot_u, J = np.unique(act_nodes, return_inverse=True)
    J = J.reshape(act_nodes.shape)
    N0 = coords.shape[0]
    pn = np.zeros((act_nodes.shape[0], nf * nz), np.int64)
    pn[:, :nf] = act_nodes
    for k in range(1, nz):
        pn[:, k * nf:(k + 1) * nf] = N0 + (k - 1) * bot_u.size + J
    vc = np.zeros((N0 + (nz - 1) * bot_u.size, 3))
    vc[:N0] = coords
    for c, el in enumerate(act_nodes):
        for k in range(nz):
            ids = pn[c, k * nf:(k + 1) * nf]
            vc[ids, 0] = coords[el, 0]
            vc[ids, 1] = coords[el, 1]
            vc[ids, 2] = coords[el, 2] + (ksi[k] + 1) / 2 * t_pzt + h / 2
    return pn, vc


def find_boundary_nodes(nodes2d, nx):
    """``find_boundary_nodes.m``: corner nodes shared with at most one other element's
    corners, edge-interior nodes on no other element's edges."""
    nodes2d = np.asarray(nodes2d, np.int64)
    mc = [0, nx - 1, nx * nx - nx, nx * nx - 1]
    e1 = list(range(1, nx - 1))
    e2 = [(k + 1) * nx - 1 for k in range(1, nx - 1)]
    e3 = list(range(nx * nx - nx + 1, nx * nx - 1))
    e4 = [k * nx for k in range(1, nx - 1)]
    edges = e1 + e2 + e3 + e4
    out = []
    for k, el in enumerate(nodes2d):
        others = np.delete(nodes2d, k, axis=0)
        for i in mc:
            if np.sum(others[:, mc] == el[i]) <= 1:
                out.append(el[i])
        for i in edges:
            if np.sum(others[:, edges] == el[i]) == 0:
                out.append(el[i])
    return np.unique(out)


def shell_actuator(nx, nz, act_nodes, coords, t_pzt, h, e, g, V0):
    """Fa [5 N] of ``main_flat_shell_multi_pzt_c.m:158-235``."""
    pn, vc = connect_pzt3d_to_plate(nx, nz, act_nodes, coords, t_pzt, h)
    fa = actuator_forces(pn, vc, e, g, nx, nz, V0, frame=0)
    bn = find_boundary_nodes(np.asarray(act_nodes), nx)
    N = coords.shape[0]
    Fa = np.zeros(5 * N)
    Fa[5 * bn] = fa[3 * bn]
    Fa[5 * bn + 2] = fa[3 * bn + 1]
    Fa[5 * bn + 1] = fa[3 * bn] * (h / 2 + t_pzt / 2)
    Fa[5 * bn + 3] = fa[3 * bn + 1] * (h / 2 + t_pzt / 2)
    return Fa


def shell_sensor_voltage(model, U, elems, t_pzt, h, e31, e32, g33):
    """``main_flat_shell_multi_pzt_c.m:618-623`` on an oracle ShellModel."""
    Npx, Npy = model.Npx, model.Npy
    I = np.concatenate([np.arange(e * model.nloc, (e + 1) * model.nloc) for e in elems])
    UX, UY, FIX, FIY = U[model.Iu], U[model.Iv], U[model.Ifix], U[model.Ifiy]
    NpxUX, NpyUX, NpxFIX, NpyFIX = Npx @ UX, Npy @ UX, Npx @ FIX, Npy @ FIX
    NpxUY, NpyUY, NpxFIY, NpyFIY = Npx @ UY, Npy @ UY, Npx @ FIY, Npy @ FIY
    i11, i12, i21, i22 = model.invJ11[I], model.invJ12[I], model.invJ21[I], model.invJ22[I]
    V = t_pzt / g33 * (e31 * (NpxUX[I] * i11 + NpyUX[I] * i12 + h / 2 * NpxFIX[I] * i11 + h / 2 * NpyFIX[I] * i12)
                       + e32 * (NpxUY[I] * i21 + NpyUY[I] * i22 + h / 2 * NpxFIY[I] * i21 + h / 2 * NpyFIY[I] * i22))
    return float(np.mean(V))


def laminate_shell_add_pzt(row, h, t_pzt, rho_pzt, Q):
    """Coefficient row (23, SHELL_COEFS order) with the transducer layer added."""
    r = np.array(row, float)
    h2, h1 = h / 2 + t_pzt / 2, h / 2
    sh = 5 / 4 * (h2 - h1 - 4 / 3 * (h2 ** 3 - h1 ** 3) / (h + t_pzt) ** 2)
    q = lambda i, j: Q[i - 1, j - 1]
    for k, (i, j) in enumerate([(1, 1), (1, 2), (1, 6), (2, 2), (2, 6), (6, 6)]):
        r[k] += q(i, j) * (h2 - h1)
        r[9 + k] += q(i, j) * (h2 ** 2 - h1 ** 2) / 2
        r[15 + k] += q(i, j) * (h2 ** 3 - h1 ** 3) / 3
    r[6] += q(4, 4) * sh
    r[7] += q(4, 5) * sh
    r[8] += q(5, 5) * sh
    r[21] += rho_pzt * t_pzt
    r[22] += rho_pzt * (h2 ** 3 - h1 ** 3) / 3
    return r
