"""Delamination model (CPU): the node split of ``split_delam_nodes_flat_shell.m`` and the
sub-laminates of ``composite_plate_c_delam.m:89-173``, oracle pins and C-ABI equality.

Pins: closed-form node counts of rectangular regions, coincident twins, rim nodes shared
with the pristine elements, the free delaminated plate keeping exactly its 6 rigid modes
(a planted mistake -- splitting the rim too -- cuts the copies loose and adds 6 more), and
the sub-laminates partitioning the whole laminate's integrals."""
import numpy as np
import pytest

import semhip
from oracle import mesh as om, shell as osh

LH = [0.25e-3] * 8
ANG = [0, 90, 0, 90, 90, 0, 90, 0]


def _mesh(nex, ney, order, lx=None, ly=None):
    lx = 0.012 * nex if lx is None else lx
    ly = 0.010 * ney if ly is None else ly
    q, p = om.rect_quad_mesh(lx, ly, nex, ney)
    return om.quad2spectral(q, p, order)


def _elem(i, j, nex):
    return j * nex + i


def test_boundary_nodes_structured():
    for order in (2, 4, 5):
        nodes, coords = _mesh(5, 4, order)
        b = om.boundary_nodes(nodes, order + 1)
        nx, ny = 5 * order + 1, 4 * order + 1
        assert b.size == 2 * (nx + ny) - 4
        x, y = coords[b, 0], coords[b, 1]
        on_edge = np.isclose(x, 0) | np.isclose(x, x.max()) | np.isclose(y, 0) | np.isclose(y, y.max())
        assert on_edge.all()
        assert np.array_equal(semhip.boundary_nodes(nodes, order + 1), b)


@pytest.mark.parametrize("a,b,order", [(2, 2, 4), (3, 1, 3), (1, 1, 5), (3, 2, 2)])
def test_split_interior_rectangle_closed_form(a, b, order):
    nex, ney = 6, 5
    nodes, coords = _mesh(nex, ney, order)
    n = order + 1
    den = [_elem(1 + i, 1 + j, nex) for j in range(b) for i in range(a)]
    bnd = om.boundary_nodes(nodes, n)
    n2, c2, du, da = om.split_delam_nodes(nodes, coords, den, n, n, bnd)
    # one new node per interior node of the region; the rim (2(a p + b p) nodes) is shared
    n_inner = (a * order - 1) * (b * order - 1)
    assert c2.shape[0] == coords.shape[0] + n_inner
    assert n2.shape[0] == nodes.shape[0] + len(den)
    assert np.array_equal(du, den) and np.array_equal(da, nodes.shape[0] + np.arange(len(den)))
    assert np.array_equal(n2[: nodes.shape[0]], nodes)
    # twins coincide; the copies use rim nodes (shared with pristine elements) or new nodes
    assert np.array_equal(c2[n2[da]], c2[n2[du]])
    new = n2[da] >= coords.shape[0]
    pristine = np.setdiff1d(np.arange(nodes.shape[0]), den)
    rim = np.intersect1d(np.unique(n2[da][~new]), np.unique(nodes[pristine]))
    assert rim.size == 2 * (a + b) * order
    assert np.array_equal(np.unique(n2[da][~new]), rim)
    assert not np.isin(n2[da][new], n2[: nodes.shape[0]]).any()
    # C-ABI equality, bit-exact
    r = semhip.split_delam(nodes, coords, den, n, bnd)
    for x, y in zip(r, (n2, c2, du, da)):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("den", [[0], [0, 1, 6], [4, 5, 10, 11], [0, 1, 2, 3, 4, 5], [7, 8, 13, 14, 15, 20]])
def test_split_regions_touching_the_boundary_match_cabi(den):
    nodes, coords = _mesh(6, 5, 4)
    r = semhip.split_delam(nodes, coords, den, 5)
    o = om.split_delam_nodes(nodes, coords, den, 5, 5, om.boundary_nodes(nodes, 5))
    for x, y in zip(r, o):
        assert np.array_equal(x, y)


def _delam_model(split_rim=False):
    nex, ney, order = 4, 3, 3
    nodes, coords = _mesh(nex, ney, order)
    n = order + 1
    den = [_elem(1, 1, nex), _elem(2, 1, nex)]
    n2, c2, du, da = om.split_delam_nodes(nodes, coords, den, n, n, om.boundary_nodes(nodes, n))
    if split_rim:
        # planted mistake: every node of the copies split, rim included
        nxt = c2.shape[0]
        old = np.unique(n2[da][n2[da] < coords.shape[0]])
        ren = {int(v): nxt + k for k, v in enumerate(old)}
        n2 = n2.copy()
        n2[da] = np.vectorize(lambda v: ren.get(int(v), int(v)))(n2[da])
        c2 = np.concatenate([c2, coords[old]], axis=0)
    Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    rows = [osh.laminate_shell(LH, ANG, 1570, **Q), osh.laminate_shell(LH, ANG, 1570, plies=range(3), **Q),
            osh.laminate_shell(LH, ANG, 1570, plies=range(3, 8), **Q)]
    coef = np.array([[r[k] for k in osh.SHELL_COEFS] for r in rows])
    em = np.zeros(n2.shape[0], np.int64)
    em[da] = 1
    em[du] = 2
    return osh.ShellModel(n2, c2, order, coef, elem_mat=em), c2


@pytest.mark.parametrize("split_rim,n_zero", [(False, 6), (True, 12)])
def test_delaminated_plate_rigid_modes(split_rim, n_zero):
    m, _ = _delam_model(split_rim)
    K = np.stack([m.internal_force(e) for e in np.eye(m.dof)], axis=1)
    assert np.abs(K - K.T).max() < 1e-12 * np.abs(K).max()
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert w[0] > -1e-12 * w[-1]
    assert np.sum(np.abs(w) < 1e-12 * w[-1]) == n_zero


def test_sublaminates_partition_the_laminate():
    Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    q = [Q[k] for k in ("q11", "q12", "q22", "q44", "q55", "q66")]
    full = semhip.laminate_shell(LH, ANG, 1570, q)
    for L in range(1, 8):
        up = semhip.laminate_shell_delam(LH, ANG, 1570, q, L, +1)
        dn = semhip.laminate_shell_delam(LH, ANG, 1570, q, L, -1)
        assert np.allclose(up + dn, full, rtol=1e-13, atol=1e-13 * np.abs(full).max())
        o_up = osh.laminate_shell(LH, ANG, 1570, plies=range(L), **Q)
        o_dn = osh.laminate_shell(LH, ANG, 1570, plies=range(L, 8), **Q)
        assert np.allclose(up, [o_up[k] for k in osh.SHELL_COEFS], rtol=1e-14, atol=0)
        assert np.allclose(dn, [o_dn[k] for k in osh.SHELL_COEFS], rtol=1e-14, atol=0)
    # symmetric layup split at the mid-plane: equal membrane terms, opposite coupling
    up = semhip.laminate_shell_delam(LH, ANG, 1570, q, 4, +1)
    dn = semhip.laminate_shell_delam(LH, ANG, 1570, q, 4, -1)
    i_a, i_b, i_d = slice(0, 6), slice(9, 15), slice(15, 21)
    assert np.allclose(up[i_a], dn[i_a], rtol=1e-14)
    assert np.allclose(up[i_b], -dn[i_b], rtol=1e-12)
    assert np.abs(up[i_b]).max() > 0
    assert np.allclose(up[i_d], dn[i_d], rtol=1e-13)
    with pytest.raises(semhip.SemError):
        semhip.laminate_shell_delam(LH, ANG, 1570, q, 8, +1)
