"""Wavefield on a uniform grid (CPU): the oracle's restatement of
``spec2meshgrid_flat_shell.m`` pinned by closed forms, and the library's host set-up
(owner search, Newton inversion) against it.

Pins: polynomial fields of degree <= p are reproduced exactly on affine meshes; the
Lagrange product basis equals the reference's inverse-Vandermonde basis; on distorted
meshes the Newton iteration maps every local point back onto its grid point (skipping the
iterations -- a planted mistake -- leaves errors of micrometres); every owner holds its
point; delaminated copies are never owners."""
import numpy as np
import pytest

import semhip
from semhip import grid as G
from oracle import mesh as om, grid as og
from oracle.common import gll


def _mesh(nex=6, ney=5, order=4, perturb=0.0, seed=0, lx=0.06, ly=0.05):
    q, p = om.rect_quad_mesh(lx, ly, nex, ney)
    if perturb:
        rng = np.random.default_rng(seed)
        inner = (p[:, 0] > 0) & (p[:, 0] < lx) & (p[:, 1] > 0) & (p[:, 1] < ly)
        p = p.copy()
        p[inner, :2] += perturb * rng.uniform(-1, 1, (inner.sum(), 2)) * min(lx / nex, ly / ney)
    return om.quad2spectral(q, p, order)


@pytest.mark.parametrize("order", [2, 4, 5])
def test_polynomial_fields_exact_on_affine_mesh(order):
    nodes, coords = _mesh(order=order)
    x, y = coords[:, 0], coords[:, 1]
    U = np.zeros((coords.shape[0], 5))
    U[:, 4] = 3 * x ** order - 2 * x * y + y ** order + 0.5   # degree <= p in each of x, y
    r = og.spec2meshgrid(nodes, coords, order + 1, [U.reshape(-1)], 31, 27, 3)
    XI, YI = np.meshgrid(r["xi"], r["yi"])
    exact = 3 * XI ** order - 2 * XI * YI + YI ** order + 0.5
    assert np.abs(r["data"][0] - exact).max() < 1e-13 * np.abs(exact).max()


def test_lagrange_basis_equals_reference_vandermonde_form():
    rng = np.random.default_rng(1)
    for n in (3, 5, 6, 8):
        ksi, _ = gll(n)
        s, t = rng.uniform(-1, 1, 11), rng.uniform(-1, 1, 11)
        Nv = og.vandermonde_shape(n, ksi, s, t)
        Ls, Lt = og.lagrange_basis(ksi, s), og.lagrange_basis(ksi, t)
        Nl = (Ls[:, None, :] * Lt[:, :, None]).reshape(-1, n * n).T
        assert np.abs(Nv - Nl).max() < 1e-12


@pytest.mark.parametrize("iterations,ok", [(3, True), (0, False)])
def test_newton_maps_back_onto_the_grid_points(iterations, ok):
    nodes, coords = _mesh(perturb=0.25, seed=3)
    n = 5
    xi, yi, xp, yp = og.grid_points(coords, 40, 35)
    owner = og.owner_elements(nodes, coords, n, nodes.shape[0], xp, yp)
    s, t, N = og.local_coordinates(nodes, coords, n, owner, xp, yp, iterations=iterations)
    en = nodes[owner]
    err = np.hypot((N * coords[en, 0]).sum(1) - xp, (N * coords[en, 1]).sum(1) - yp)
    assert (err.max() < 1e-12) == ok
    if ok:
        assert np.abs(s).max() < 1 + 1e-9 and np.abs(t).max() < 1 + 1e-9


def test_cabi_plan_matches_oracle_on_a_distorted_mesh():
    nodes, coords = _mesh(perturb=0.25, seed=3)
    n, Nx, Ny = 5, 40, 35
    xi, yi, xp, yp = og.grid_points(coords, Nx, Ny)
    owner = og.owner_elements(nodes, coords, n, nodes.shape[0], xp, yp)
    s, t, _ = og.local_coordinates(nodes, coords, n, owner, xp, yp)
    g = G.grid_plan(nodes, coords, n, Nx, Ny)
    assert np.array_equal(g["xi"], xi) and np.array_equal(g["yi"], yi)
    assert np.array_equal(g["owner"], owner)
    assert np.abs(g["s"] - s).max() < 1e-12 and np.abs(g["t"] - t).max() < 1e-12
    # every owner holds its point (corner polygon, counter-clockwise)
    mc = [0, n - 1, n * n - 1, n * n - n]
    X, Y = coords[nodes[owner][:, mc], 0], coords[nodes[owner][:, mc], 1]
    for q in range(4):
        q1 = (q + 1) % 4
        cross = (X[:, q1] - X[:, q]) * (yp - Y[:, q]) - (Y[:, q1] - Y[:, q]) * (xp - X[:, q])
        assert (cross > -1e-15).all()


def test_delaminated_copies_are_not_searched():
    nodes, coords = _mesh(order=4)
    n2, c2, du, da = semhip.split_delam(nodes, coords, [8, 9, 14, 15], 5)
    g = G.grid_plan(n2, c2, 5, 30, 30, n_search=nodes.shape[0])
    assert g["owner"].max() < nodes.shape[0]
    assert np.isin(du, g["owner"]).all()
    xi, yi, xp, yp = og.grid_points(c2, 30, 30)
    assert np.array_equal(g["owner"], og.owner_elements(n2, c2, 5, nodes.shape[0], xp, yp))


def test_variable_names_follow_the_reference():
    assert G.variable_name("displacement", 6) == "Ux_top"
    assert G.variable_name("velocity", 4) == "VFix"
    assert G.variable_name("acceleration", 9) == "total_accelerations"
    assert G.variable_name("velocity", 8) == "velocities_in_plane"


def test_grid_point_outside_every_element_is_an_error():
    """An L-shaped mesh leaves grid points of its bounding box outside every element: the
    reference would index an empty candidate list; the library reports an error."""
    nodes, coords = _mesh(nex=4, ney=4, order=2, lx=0.04, ly=0.04)
    cen = coords[nodes].mean(axis=1)
    keep = ~((cen[:, 0] > 0.02) & (cen[:, 1] > 0.02))          # drop the top-right quarter
    with pytest.raises(semhip.SemError):
        G.grid_plan(nodes[keep], coords, 3, 9, 9)
    assert (G.grid_plan(nodes, coords, 3, 9, 9)["owner"] >= 0).all()   # the full square is fine
