"""PZT transducer set-up: the library's C-ABI (host code, no GPU) against the oracle, and
both against what the physics fixes.

* strain-charge -> stress-charge conversion reproduces PZT-5A's published constants;
* an axis-aligned transducer block under V0: the oracle's induced potential is linear
  through the thickness, the forces are self-equilibrated and the resultant on an edge
  face is e31 V0 L; the open-circuit voltage of a uniform strain eps_x is -e31 eps t / g33;
* the flat-shell actuator loads exactly the footprint's perimeter nodes.
"""
import numpy as np
import pytest

from semhip import pzt
from oracle import mesh as om
from oracle import pzt as op
from oracle import shell as osh
from cases import rel_l2

EPS0 = 8.854187817e-12


def _pzt5a_strain_charge():
    """PZT-5A strain-charge data (columns ex ey ez gxz gyz gxy)."""
    s11, s12, s13, s33, s44 = 16.4e-12, -5.74e-12, -7.22e-12, 18.8e-12, 47.5e-12
    s66 = 2 * (s11 - s12)
    S = np.array([[s11, s12, s13, 0, 0, 0], [s12, s11, s13, 0, 0, 0], [s13, s13, s33, 0, 0, 0],
                  [0, 0, 0, s44, 0, 0], [0, 0, 0, 0, s44, 0], [0, 0, 0, 0, 0, s66]])
    d = np.zeros((3, 6))
    d[0, 3] = d[1, 4] = 584e-12
    d[2, :3] = [-171e-12, -171e-12, 374e-12]
    epsT = np.diag([1730 * EPS0, 1730 * EPS0, 1700 * EPS0])
    return S, d, epsT


def _block(nex=2, ney=2, order=3, order_z=2, L=0.01, t=0.5e-3):
    """Stand-alone transducer block [0, L]^2 x [0, t] of hexahedra (bottom face first)."""
    q, p = om.rect_quad_mesh(L, L, nex, ney)
    n2, c2 = om.quad2spectral(q, p, order)
    n3, c3 = om.extrude_mesh(n2, c2, order + 1, order + 1, order_z + 1, [t], 1)
    return n3, c3, order + 1, order_z + 1


def test_pzt_const_known_answer():
    S, d, epsT = _pzt5a_strain_charge()
    Q, e, g, _ = pzt.pzt_const(S, d, epsT, "solid")
    assert abs(Q[0, 0] / 121e9 - 1) < 0.02
    assert abs(Q[2, 2] / 111e9 - 1) < 0.02
    assert abs(e[2, 0] / -5.4 - 1) < 0.02
    assert abs(e[2, 2] / 15.8 - 1) < 0.02
    assert abs(g[2, 2] / 7.35e-9 - 1) < 0.02


@pytest.mark.parametrize("variant,theta", [("solid", 0.0), ("shell", 0.0), ("shell", 30.0)])
def test_pzt_const_matches_oracle(variant, theta):
    S, d, epsT = _pzt5a_strain_charge()
    Q, e, g, x = pzt.pzt_const(S, d, epsT, variant, theta)
    if variant == "solid":
        Qo, eo, go = op.pzt_const_solid(S, d, epsT)
    else:
        Qo, eo, go, e31, e32, g33 = op.pzt_const_shell(S, d, epsT, theta)
        assert np.allclose(x, [e31, e32, g33], rtol=1e-12)
    assert rel_l2(Q, Qo) < 1e-12 and rel_l2(e, eo) < 1e-12 and rel_l2(g, go) < 1e-12


@pytest.mark.parametrize("frame", [0, 1])
def test_actuator_matches_oracle_and_closed_form(frame):
    n3, c3, nx, nz = _block()
    Q, e, g, _ = pzt.pzt5a()
    V0, L, t = 10.0, 0.01, 0.5e-3
    fa = pzt.actuator(n3, c3, nx, nz, e, g, V0, frame)
    fo, phi = op.actuator_forces(n3, c3, e, g, nx, nz, V0, frame, return_potential=True)
    assert rel_l2(fa, fo) < 1e-10
    # linear potential through the thickness
    assert np.allclose(phi, V0 * c3[:, 2] / t, rtol=0, atol=1e-9 * V0)
    f = fa.reshape(-1, 3)
    assert np.abs(f.sum(axis=0)).max() < 1e-9 * np.abs(f).max()
    xmax = np.isclose(c3[:, 0], L)
    xmin = np.isclose(c3[:, 0], 0.0)
    r = f[xmax, 0].sum()
    assert abs(abs(r) - abs(e[2, 0]) * V0 * L) < 1e-9 * abs(r)
    assert abs(f[xmin, 0].sum() + r) < 1e-9 * abs(r)


@pytest.mark.parametrize("frame", [0, 1])
def test_sensor_matches_oracle_and_closed_form(frame):
    n3, c3, nx, nz = _block()
    Q, e, g, _ = pzt.pzt5a()
    t = 0.5e-3
    w = pzt.sensor(n3, c3, nx, nz, e, g, frame)
    rng = np.random.default_rng(0)
    for _ in range(2):
        U = rng.standard_normal(3 * c3.shape[0]) * 1e-6
        v = op.sensor_voltage(n3, c3, e, g, nx, nz, U, frame)
        assert abs(w @ U - v) < 1e-10 * abs(v)
    assert rel_l2(w, op.sensor_weights(n3, c3, e, g, nx, nz, frame)) < 1e-10
    eps = 1e-5
    U = np.zeros((c3.shape[0], 3))
    U[:, 0] = eps * c3[:, 0]
    ref = -e[2, 0] * eps * t / g[2, 2]
    assert abs(w @ U.reshape(-1) - ref) < 1e-9 * abs(ref)


def test_surface_transducer_condenses_columns():
    """Forces of a surface transducer keep their resultant when carried down to the plate,
    and its voltage weights see only plate DOFs under the footprint."""
    from cases import solid_pair
    _, inp = solid_pair(nex=3, ney=3, nlay=1, order=2, order_z=2, lx=0.03, ly=0.03)
    nodes, coords = inp["nodes"], inp["coords"]
    Q, e, g, _ = pzt.pzt5a()
    el = pzt.top_elements(nodes, coords, (0.015, 0.015), 0.006, 3, 3)
    assert len(el) == 1
    pn, X, col = pzt.surface_transducer(nodes, coords, 3, 3, el, 0.2e-3)
    full = pzt.actuator(pn, X, 3, 3, e, g, 1.0)
    dofs, amp = pzt.surface_actuator(nodes, coords, 3, 3, el, 0.2e-3, e, g)
    f = np.zeros(coords.shape[0] * 3)
    f[dofs] = amp
    assert np.allclose(f.reshape(-1, 3).sum(0), full.reshape(-1, 3).sum(0), atol=1e-12 * np.abs(full).max())
    top = np.isclose(coords[:, 2], coords[:, 2].max())
    assert np.all(top[dofs // 3])
    sd, sw = pzt.surface_sensor(nodes, coords, 3, 3, el, 0.2e-3, e, g)
    assert np.all(top[sd // 3])


def test_shell_actuator_matches_oracle_and_loads_perimeter():
    q, p = om.rect_quad_mesh(0.04, 0.04, 4, 4)
    nodes, coords = om.quad2spectral(q, p, 3)
    S, d, epsT = _pzt5a_strain_charge()
    Q, e, g, x = pzt.pzt_const(S, d, epsT, "shell")
    cen = coords[nodes].mean(axis=1)
    act = np.nonzero((np.abs(cen[:, 0] - 0.02) < 0.011) & (np.abs(cen[:, 1] - 0.02) < 0.011))[0]
    assert len(act) == 4
    h, t = 2e-3, 0.5e-3
    dofs, amp = pzt.shell_actuator(nodes[act], coords, 4, 3, t, h, e, g, 10.0)
    fa = np.zeros(5 * coords.shape[0])
    fa[dofs] = amp
    fo = op.shell_actuator(4, 3, nodes[act], coords, t, h, e, g, 10.0)
    assert rel_l2(fa, fo) < 1e-10
    loaded = np.unique(dofs // 5)
    xs, ys = coords[loaded, 0], coords[loaded, 1]
    on_perimeter = np.isclose(xs, 0.01) | np.isclose(xs, 0.03) | np.isclose(ys, 0.01) | np.isclose(ys, 0.03)
    assert np.all(on_perimeter)
    foot = np.unique(nodes[act])
    fx, fy = coords[foot, 0], coords[foot, 1]
    perim = foot[np.isclose(fx, 0.01) | np.isclose(fx, 0.03) | np.isclose(fy, 0.01) | np.isclose(fy, 0.03)]
    assert set(op.find_boundary_nodes(nodes[act], 4)) == set(perim)


def test_shell_laminate_pzt_layer_matches_oracle():
    Qs = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    lam = osh.laminate_shell([0.25e-3] * 8, [0, 90, 0, 90, 90, 0, 90, 0], 1570, **Qs)
    row = np.array([lam[k] for k in osh.SHELL_COEFS])
    S, d, epsT = _pzt5a_strain_charge()
    Q, e, g, x = pzt.pzt_const(S, d, epsT, "shell")
    out = pzt.laminate_shell_add_pzt(row, 2e-3, 0.5e-3, 7750, Q)
    ref = op.laminate_shell_add_pzt(row, 2e-3, 0.5e-3, 7750, Q)
    assert rel_l2(out, ref) < 1e-14
    assert out[21] == pytest.approx(row[21] + 7750 * 0.5e-3, rel=1e-14)


def test_solid_plate_250x250_setup():
    """The 250 x 250 plate set-up (on a coarse grid): PZT actuator forces at the centre, 64 surface
    w-displacement sensors on the 8 x 8 grid, 20 snapshots over the run."""
    from semhip import engine as E
    c = E.solid_plate_250x250_pzt(nex=20, ney=20, order=3, order_z=3, nft=600, pzt_size=0.1)  # 2 x 2 elements of 50 mm
    prep = E.prepare_solid(c)
    assert len(prep["outputs"]) == 64
    top = np.isclose(prep["coords"][:, 2], prep["coords"][:, 2].max())
    outs = np.asarray(prep["outputs"])
    assert (outs % 3 == 2).all() and top[outs // 3].all()
    xy = prep["coords"][outs // 3, :2]
    assert np.allclose(xy, np.asarray(E.SURFACE_GRID_64), atol=0.5 / 20 * 0.5 + 1e-12)
    assert len(prep["frame_no"]) == 20 and prep["frame_no"][-1] == 600
    f = np.zeros(prep["coords"].shape[0] * 3)
    np.add.at(f, prep["ex_dofs"], prep["ex_amp"])
    assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-9 * np.abs(f).max()   # self-equilibrated
    near = (np.abs(prep["coords"][:, 0] - 0.5) < 0.05 + 1e-9) & (np.abs(prep["coords"][:, 1] - 0.5) < 0.05 + 1e-9)
    assert np.abs(f.reshape(-1, 3)[~near]).max() == 0.0


def _asymmetric_strain_charge():
    """Artificial strain-charge data whose shear terms all differ (columns ex ey ez, then the
    three shears in the input's columns 4, 5, 6): a swap of any two shear columns changes the
    coupling, which PZT-5A's symmetric constants (d15 = d24, s44 = s55) would hide."""
    S = np.diag([16.4e-12, 17.1e-12, 18.8e-12, 41.0e-12, 47.5e-12, 53.0e-12])
    S[0, 1] = S[1, 0] = -5.74e-12
    S[0, 2] = S[2, 0] = -7.22e-12
    S[1, 2] = S[2, 1] = -6.9e-12
    d = np.zeros((3, 6))
    d[0, 3:] = [584e-12, 211e-12, 97e-12]
    d[1, 3:] = [143e-12, 611e-12, 59e-12]
    d[2, :] = [-171e-12, -163e-12, 374e-12, 31e-12, 17e-12, 23e-12]
    epsT = np.diag([1730 * EPS0, 1650 * EPS0, 1700 * EPS0])
    return S, d, epsT


# solid3D/pzt_const.m:20-22: the rearranged e has columns (1-based) 1, 2, 3, 6, 5, 4 of
# ep = d inv(S); coupl3D_v2.m:121-133 orders the strain rows ex, ey, ez, gxy (du/dy+dv/dx),
# gyz (dw/dy+dv/dz), gxz (dw/dx+du/dz)
EP_COLUMN_OF_STRAIN_ROW = {"gxy": 6, "gyz": 5, "gxz": 4}


@pytest.mark.parametrize("frame", [0, 1])
@pytest.mark.parametrize("shear,field", [("gyz", 0), ("gyz", 1), ("gxz", 0), ("gxz", 1), ("gxy", 0), ("gxy", 2)])
def test_shear_coupling_against_closed_form(frame, shear, field):
    """One axis-aligned brick, a linear displacement field whose only strain is one shear
    gamma and a linear potential phi = x_field: u^T kufi phi = V gamma ep[field, c], the
    integral of the constant electric displacement, with c the column of ep = d inv(S) that
    the reference pairs with that strain row.  A swapped pair of shear columns (e15/e24)
    gives another entry of ep (all differ here) and fails."""
    S, d, epsT = _asymmetric_strain_charge()
    Qo, eo, go = op.pzt_const_solid(S, d, epsT)
    ep = d @ np.linalg.inv(S)
    n3, c3, nx, nz = _block(nex=1, ney=1, order=3, order_z=2, L=0.01, t=0.5e-3)
    X = c3[n3[0]]
    kufi, _ = op.coupl3d(eo, go, X, nx, nz, frame)
    gam = 1e-4
    U = np.zeros_like(X)
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    if shear == "gyz":
        U[:, 2] = gam * y
    elif shear == "gxz":
        U[:, 2] = gam * x
    else:
        U[:, 0] = gam * y
    phi = X[:, field]
    V = 0.01 * 0.01 * 0.5e-3
    got = U.reshape(-1) @ kufi @ phi
    ref = V * gam * ep[field, EP_COLUMN_OF_STRAIN_ROW[shear] - 1]
    assert abs(got - ref) < 1e-9 * np.abs(ep).max() * V * gam, (got, ref)


@pytest.mark.parametrize("frame", [0, 1])
def test_asymmetric_shear_constants_library_matches_oracle(frame):
    """The library's constants, actuator forces and sensor weights with shear constants that
    all differ, against the oracle (pinned above): the random displacements of the sensor
    check exercise every shear term of the coupling."""
    S, d, epsT = _asymmetric_strain_charge()
    Q, e, g, _ = pzt.pzt_const(S, d, epsT, "solid")
    Qo, eo, go = op.pzt_const_solid(S, d, epsT)
    assert rel_l2(Q, Qo) < 1e-12 and rel_l2(e, eo) < 1e-12 and rel_l2(g, go) < 1e-12
    n3, c3, nx, nz = _block()
    rng = np.random.default_rng(5)
    c3 = c3 + rng.uniform(-2e-5, 2e-5, c3.shape) * np.array([1, 1, 0.02])  # not axis-aligned
    w = pzt.sensor(n3, c3, nx, nz, e, g, frame)
    assert rel_l2(w, op.sensor_weights(n3, c3, eo, go, nx, nz, frame)) < 1e-10
    fa = pzt.actuator(n3, c3, nx, nz, e, g, 10.0, frame)
    assert rel_l2(fa, op.actuator_forces(n3, c3, eo, go, nx, nz, 10.0, frame)) < 1e-10
