"""Mesh front end of the oracle (TEST INFRASTRUCTURE).

* ``rect_quad_mesh``  -- a structured rectangle of bilinear quads, corner nodes numbered
                         row-major (x fastest), quads counter-clockwise.  This stands in for
                         gmsh's transfinite output (``flat_shell/geo/rect_50x50.geo`` via
                         ``automesh_rectangle.m:47-53``); gmsh is absent, so the corner
                         numbering is this build's own convention (parity of what follows is
                         defined on it).
* ``change_turn_quad``-- orientation check of ``mesher/change_turn_quad.m`` (flip quads with a
                         non-positive Jacobian determinant).
* ``quad2spectral``   -- ``mesher/quad2spectral_Fiborek.m:1-173``: corner nodes keep their
                         numbers, edge-interior nodes are numbered by ``unique(...,'rows')``
                         (lexicographic x, y, z) of their coordinates, element-interior nodes
                         follow element by element; local order xi fastest.
                         Pinned by: the committed gmsh fixture ``m1_rand_2`` (conformity and
                         geometry invariants), exact grid coordinates on structured meshes,
                         every node used, shared edges shared.  The exact numbering of the
                         edge nodes is "parity unpinned" beyond those invariants: no reference
                         output exists for it (gmsh + MATLAB cannot run here).
* ``extrude_mesh``    -- ``solid3D/mesher3D/extrude_mesh.m:1-85`` including its ``hlay = dz``
                         update (``:84``), which is exact only for Nz <= 2.
"""
from __future__ import annotations

import numpy as np

from .common import gll


def rect_quad_mesh(lx, ly, nex, ney, x0=0.0, y0=0.0):
    """Structured quads over [x0, x0+lx] x [y0, y0+ly]; returns (quads[fen,4] 0-based, pos[N,3])."""
    xs = x0 + lx * np.arange(nex + 1) / nex
    ys = y0 + ly * np.arange(ney + 1) / ney
    X, Y = np.meshgrid(xs, ys)  # [ney+1, nex+1]
    pos = np.stack([X.reshape(-1), Y.reshape(-1), np.zeros(X.size)], axis=1)
    j, i = np.meshgrid(np.arange(ney), np.arange(nex), indexing="ij")
    n1 = (j * (nex + 1) + i).reshape(-1)
    quads = np.stack([n1, n1 + 1, n1 + nex + 2, n1 + nex + 1], axis=1)
    return quads.astype(np.int64), pos


def change_turn_quad(quads, pos):
    """``change_turn_quad.m``: flip the node order [1,4,3,2] of quads whose bilinear
    Jacobian determinant is <= 0 at a corner (the reference evaluates it at the 4 corners)."""
    quads = quads.copy()
    p = pos[quads]  # [fen,4,3]
    # corner Jacobian determinant = cross(next - a, prev - a) at each corner a
    flips = np.zeros(len(quads), dtype=bool)
    for a, b, c in ((0, 1, 3), (1, 2, 0), (2, 3, 1), (3, 0, 2)):
        e1 = p[:, b, :2] - p[:, a, :2]
        e2 = p[:, c, :2] - p[:, a, :2]
        det = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
        flips |= det <= 0
    quads[flips] = quads[flips][:, [0, 3, 2, 1]]
    return quads


def _matlab_round(v):
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def quad2spectral(quads, pos, order):
    """``quad2spectral_Fiborek.m:1-173`` restated with 0-based ids.

    Returns (nodes[fen, n*n], coords[N, 3]) with n = order+1.
    """
    quads = np.asarray(quads, dtype=np.int64)
    pos = np.asarray(pos, dtype=np.float64)
    n = order + 1
    if n < 3 or n > 10:
        raise ValueError("need 3..10 nodes per edge")
    fen = quads.shape[0]
    ni = n - 2
    ksi, _ = gll(n)
    kint = ksi[1:n - 1]
    # edgegeometry (:155-170): per element, 4 edges 1->2, 2->3, 3->4, 4->1
    cyc = quads[:, [0, 1, 2, 3, 0]]
    coords_e = []
    for d in range(3):
        nC = pos[cyc, d]  # [fen, 5]
        mid = (nC[:, 1:5] + nC[:, 0:4]) / 2
        mid = _matlab_round(mid.reshape(-1) * 1e6) * 1e-6  # element-major, edge-minor
        ln = ((nC[:, 1:5] - nC[:, 0:4]) / 2).reshape(-1)
        coords_e.append((mid[:, None] + ln[:, None] * kint[None, :]).reshape(-1))
    E = np.stack(coords_e, axis=1)  # [(fen*4*ni), 3], order (e, edge, k)
    # unique(...,'first','rows') (:41): sorted unique rows, n1 = index per row
    uniq, inv = np.unique(E, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    el_edge = inv.reshape(fen, 4 * ni)  # 0-based into uniq
    n_edge = uniq.shape[0]
    # element-interior nodes (:46-51), consecutive per element
    el_int = (np.arange(fen * ni * ni).reshape(fen, ni * ni)) + n_edge
    # interior coordinates by line intersection (:52-101)
    xe, ye = uniq[:, 0], uniq[:, 1]
    e1 = el_edge[:, 0:ni]
    e2 = el_edge[:, ni:2 * ni]
    e3 = el_edge[:, 3 * ni - 1:2 * ni - 1:-1] if ni > 0 else el_edge[:, 0:0]
    e4 = el_edge[:, 4 * ni - 1:3 * ni - 1:-1] if ni > 0 else el_edge[:, 0:0]
    A13 = ye[e3] - ye[e1]
    B13 = xe[e1] - xe[e3]
    C13 = (ye[e1] - ye[e3]) * xe[e1] + (xe[e3] - xe[e1]) * ye[e1]
    A42 = ye[e2] - ye[e4]
    B42 = xe[e4] - xe[e2]
    C42 = (ye[e4] - ye[e2]) * xe[e4] + (xe[e2] - xe[e4]) * ye[e4]
    # node (p over 13-lines, q over 42-lines); stored q-major, p-minor (:95-100 column order)
    a13, b13, c13 = A13[:, None, :], B13[:, None, :], C13[:, None, :]  # [fen,1,p]
    a42, b42, c42 = A42[:, :, None], B42[:, :, None], C42[:, :, None]  # [fen,q,1]
    W = a13 * b42 - b13 * a42
    Wx = b13 * c42 - c13 * b42
    Wy = c13 * a42 - a13 * c42
    Xi = (Wx / W).reshape(-1)
    Yi = (Wy / W).reshape(-1)
    zval = np.unique(uniq[:, 2]) if n_edge else np.zeros(1)
    if zval.size != 1:
        raise ValueError("quad2spectral expects a flat (constant z) mesh")
    Zi = np.full(Xi.shape, zval[0])
    nfem = int(quads.max()) + 1
    el_sem_int = np.concatenate([el_edge, el_int], axis=1) + nfem
    sem = np.concatenate([quads, el_sem_int], axis=1)  # [fen, 4 + 4ni + ni^2]
    coords = np.concatenate([pos[:nfem], uniq, np.stack([Xi, Yi, Zi], axis=1)], axis=0)
    # local (xi-fastest) reordering, elementNodes_pl3 (:104-118)
    cols = []
    cols += [0] + [4 + k for k in range(ni)] + [1]
    for r in range(ni):
        cols += [4 + 3 * ni + (ni - 1 - r)] + [4 + 4 * ni + r * ni + p for p in range(ni)] + [4 + ni + r]
    cols += [3] + [4 + 2 * ni + k for k in range(ni - 1, -1, -1)] + [2]
    nodes = sem[:, cols]
    return nodes, coords


def extrude_mesh(nodes2d, coords2d, nx, ny, nz, h, Nz):
    """``solid3D/mesher3D/extrude_mesh.m:1-85`` (0-based ids)."""
    nodes2d = np.asarray(nodes2d, dtype=np.int64)
    fen = nodes2d.shape[0]
    ksi, _ = gll(nz)
    dofl = int(nodes2d.max()) + 1
    nodes3d = np.zeros((Nz * fen, nx * ny * nz), dtype=np.int64)
    for i in range(Nz):
        for k in range(nz):
            lev = i * (nz - 1) + k
            nodes3d[i * fen:(i + 1) * fen, k * nx * ny:(k + 1) * nx * ny] = nodes2d + dofl * lev
    ntot = int(nodes3d.max()) + 1
    coords3d = np.zeros((ntot, 3))
    hlay = 0.0
    for iz in range(Nz):
        dz = h[iz]
        zz = dz * (ksi + 1) / 2 + hlay
        for k in range(nz):
            ids = nodes3d[iz * fen:(iz + 1) * fen, k * nx * ny:(k + 1) * nx * ny]
            coords3d[ids, 0] = coords2d[nodes2d, 0]
            coords3d[ids, 1] = coords2d[nodes2d, 1]
            coords3d[ids, 2] = zz[k]
        hlay = dz  # quirk of :84 kept
    return nodes3d, coords3d
