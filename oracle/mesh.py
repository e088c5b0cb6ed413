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
* ``boundary_nodes``  -- third output of ``quad2spectral_Fiborek.m:153-156``: the nodes of the
                         element edges that belong to one element only.
* ``split_delam_nodes`` -- ``mesher/split_delam_nodes_flat_shell.m:1-276``: the delaminated
                         elements are duplicated (the copies lie above the split interface);
                         the copies keep the nodes on the delamination's rim and get new,
                         coincident nodes everywhere else.  Pinned by: the rim and interior
                         counts of rectangular regions (closed form), coincident coordinates,
                         and a free delaminated plate keeping exactly the 6 rigid modes (a
                         split that cut the copies loose would add modes).
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


def boundary_nodes(nodes, n):
    """``quad2spectral_Fiborek.m:153-156``: sorted nodes of the edges owned by one element
    (an edge is identified by its two corner nodes, equivalent to the reference's unique
    edge midpoints on a conforming mesh)."""
    nodes = np.asarray(nodes, dtype=np.int64)
    edges = ([(0, k) for k in range(n)], [(k, n - 1) for k in range(n)],
             [(n - 1, k) for k in range(n)], [(k, 0) for k in range(n)])  # (row j, col i)
    count = {}
    for e in range(nodes.shape[0]):
        for ed in edges:
            a, b = nodes[e, ed[0][0] * n + ed[0][1]], nodes[e, ed[-1][0] * n + ed[-1][1]]
            key = (min(a, b), max(a, b))
            count[key] = count.get(key, 0) + 1
    out = set()
    for e in range(nodes.shape[0]):
        for ed in edges:
            a, b = nodes[e, ed[0][0] * n + ed[0][1]], nodes[e, ed[-1][0] * n + ed[-1][1]]
            if count[(min(a, b), max(a, b))] == 1:
                out.update(int(nodes[e, j * n + i]) for j, i in ed)
    return np.array(sorted(out), dtype=np.int64)


def split_delam_nodes(nodes, coords, den, nx, ny, boundary):
    """``split_delam_nodes_flat_shell.m:1-276`` with 0-based ids, a loop-for-loop
    restatement.  Returns (nodes, coords, den_under, den_above)."""
    nodes = np.asarray(nodes, dtype=np.int64)
    coords = np.asarray(coords, dtype=np.float64)
    den = [int(x) for x in den]
    fen = nodes.shape[0]
    mc = [0, nx - 1, nx * ny - nx, nx * ny - 1]                      # :40
    edge1 = list(range(1, nx - 1))                                    # :42
    edge2 = [(k + 1) * nx + nx - 1 for k in range(ny - 2)]            # :43-45
    edge3 = list(range(nx * ny - nx + 1, nx * ny - 1))                # :46
    edge4 = [(k + 1) * nx for k in range(ny - 2)]                     # :47-49
    edges = edge1 + edge2 + edge3 + edge4
    den_under = list(den)
    den_above = [fen + k for k in range(len(den))]                    # :51-52
    delamnodes = np.unique(nodes[den, :])                             # :54
    nodemax = int(nodes.max())
    nm = nodemax
    nodes = np.concatenate([nodes, nodes[den, :]], axis=0)            # :61
    outer, free_corner = [], []
    for k in range(len(den)):
        common = nodes[den_above[k]]
        others = [j for j in range(len(den)) if j != k]
        for i in range(4):                                            # corners :68-88
            cn = common[mc[i]]
            nb = sum(1 for j in others for jc in range(4) if nodes[den_above[j], mc[jc]] == cn)
            if nb in (0, 1):
                outer.append(cn)
            if nb == 0:
                free_corner.append(cn)
        for ed in (edge1, edge2, edge3, edge4):                       # edges :90-236
            for loc in ed:
                en = common[loc]
                nb = sum(1 for j in others for jc in edges if nodes[den_above[j], jc] == en)
                if nb == 0:
                    outer.append(en)
    special = np.intersect1d(np.array(free_corner, dtype=np.int64), np.asarray(boundary, dtype=np.int64))
    if special.size == 3:                                             # :240-243
        special = special[1:3]
    outer = np.setdiff1d(np.unique(np.array(outer, dtype=np.int64)), np.asarray(boundary, dtype=np.int64))
    outer = np.concatenate([outer, special])
    inner = np.setdiff1d(delamnodes, outer)                           # :247
    for nd in inner:                                                  # :249-260
        nodemax += 1
        for i in range(len(den)):
            row = nodes[den_above[i]]
            row[row == nd] = nodemax
    coords_new = np.zeros((nodemax + 1, 3))                           # :263-266
    coords_new[:nm + 1] = coords[:nm + 1]
    coords_new[nodes[den_above, :].reshape(-1)] = coords_new[nodes[den_under, :].reshape(-1)]
    return nodes, coords_new, np.array(den_under), np.array(den_above)
