"""GPU runs with PZT transducers against the oracle (FP64 1e-10, FP32 form 1e-3 relative L2).

* flat shell (main_flat_shell_multi_pzt_c.m): transducers laminated into the plate, the
  actuator's equivalent forces and moments, every patch's voltage per step;
* 3D solid with surface transducers: the actuator's forces and the open-circuit voltage
  functionals carried to the plate surface.
"""
import numpy as np
import pytest

import semhip
from semhip import engine as E
from semhip import pzt
from cases import rel_l2
from oracle import adjoint as oa
from oracle import pzt as op
from oracle import shell as osh
from oracle import solid as oso

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel,precision,tol", [(None, "fp64", 1e-10), ((1, 8), "fp64", 1e-10),
                                                   ((1, 8), "fp32_split", 1e-3), ((2, 16), "fp32_split", 1e-3)])
def test_flat_shell_multi_pzt_matches_oracle(kernel, precision, tol):
    case = E.shell_plate_40x40(lx=0.1, ly=0.1, nex=10, ney=10, nft=300, n_frames=0, transducers="pzt",
                     pzt_xy=((0.05, 0.05), (0.08, 0.05), (0.03, 0.07)), pzt_size=0.02, actuator_no=0, V0=10.0)
    s, prep = E.flat_shell_solver(case, kernel=kernel)
    r = s.run(case.dt, 1, case.nft, precision)
    s.close()
    h = prep["h"]
    S, d, epsT, rho_p = pzt.pzt5a_strain_charge()
    Qo, eo, go, e31, e32, g33 = op.pzt_const_shell(S, d, epsT, 0.0)
    row = np.asarray(prep["coef"][0])
    coef = np.stack([row, op.laminate_shell_add_pzt(row, h, case.pzt_t, rho_p, Qo)])
    model = osh.ShellModel(prep["nodes"], prep["coords"], case.order, coef, elem_mat=prep["elem_mat"])
    f = op.shell_actuator(case.order + 1, case.pzt_nz, prep["nodes"][prep["pzt_elems"][0]], prep["coords"],
                          case.pzt_t, h, eo, go, case.V0)
    states = oa.forward_states(model, case.dt, prep["st"], f, 1, case.nft)
    ref = np.array([[op.shell_sensor_voltage(model, states[nn], el, case.pzt_t, h, e31, e32, g33)
                     for el in prep["pzt_elems"]] for nn in range(1, case.nft + 1)])
    assert np.abs(ref[:, 1]).max() > 0
    e = rel_l2(r["traces"][:case.nft], ref)
    assert e < tol, e


@pytest.mark.parametrize("kernel", [(0, 4), (1, 4), (1, 8), (3, 4), (3, 8)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_solid_surface_pzt_matches_oracle(kernel, precision, tol):
    """The 3D plate with surface transducers (the bench's workload at a small size) on every
    solid element kernel, in both forms."""
    case = E.solid_plate_100x100_pzt(lx=0.04, ly=0.04, nex=8, ney=8, order=3, order_z=3, nft=250, force_xy=(0.02, 0.02),
                     sensors=((0.03, 0.02), (0.01, 0.03)))
    s, prep = E.solid_solver(case, kernel=kernel)
    r = s.run(case.dt, 2, case.nft, precision)
    s.close()
    nodes, coords = prep["nodes"], prep["coords"]
    n, nz = case.order + 1, case.order_z + 1
    Q, e, g, _ = pzt.pzt5a()

    def fp(xy):
        return pzt.top_elements(nodes, coords, xy, case.pzt_size / 2, n, nz)

    def condensed(v, col):
        out = np.zeros(3 * coords.shape[0])
        np.add.at(out.reshape(-1, 3), col, v.reshape(-1, 3))
        return out

    pn, X, col = pzt.surface_transducer(nodes, coords, n, nz, fp(case.force_xy), case.pzt_t)
    # the oracle's actuator on the same transducer mesh, carried down its columns
    f = condensed(op.actuator_forces(pn, X, e, g, n, nz, case.V0, frame=1), col)
    model = oso.SolidModel(nodes, coords, n, nz, prep["coef"])
    states = oa.forward_states(model, case.dt, prep["st"], f, 2, case.nft)
    W = []
    for xy in case.sensors:
        pn_s, X_s, col_s = pzt.surface_transducer(nodes, coords, n, nz, fp(xy), case.pzt_t)
        W.append(condensed(op.sensor_weights(pn_s, X_s, e, g, n, nz, frame=1), col_s))
    # disp after the update at step nn = u_{nn+1} = states[nn] (states[0] = u_1)
    ref = states[2:case.nft + 1] @ np.array(W).T
    assert np.abs(ref).max() > 0
    e_ = rel_l2(r["traces"][1:case.nft], ref)
    assert e_ < tol, e_


@pytest.mark.parametrize("kernel", [(0, 4), (1, 8), (3, 4), (3, 8)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_solid_meshed_pzt_reference_layout_matches_oracle(kernel, precision, tol):
    """The reference's 3D layout (main_spec3Dcuda_global_..._mat.m:198-388, 499-537, 1054-1068):
    the actuator and two sensors are PZT-5A hexahedra of the mesh (order 4, as the 100 x 100 x 2
    workload) on the plate's top faces, with the PZT material row; the actuator's forces and
    the sensors' open-circuit voltages come from their own electrical systems.  The thin
    transducer elements set the time step: 4e-9 s, half the critical step of this mesh
    (power iteration on M^-1 K gives 7.9e-9 s)."""
    case = E.solid_plate_100x100_pzt(lx=0.04, ly=0.04, nex=8, ney=8, nft=800, dt=4e-9, f_2=500e3,
                                     force_xy=(0.02, 0.02), sensors=((0.03, 0.02), (0.01, 0.03)),
                                     pzt_layout="meshed")
    s, prep = E.solid_solver(case, kernel=kernel)
    r = s.run(case.dt, 2, case.nft, precision)
    s.close()
    nodes, coords = prep["nodes"], prep["coords"]
    n, nz = case.order + 1, case.order_z + 1
    Q, e, g, rho_p = pzt.pzt5a()
    n_plate = int((prep["elem_mat"] == 0).sum())
    assert (prep["elem_mat"][n_plate:] == 1).all() and nodes.shape[0] - n_plate == 12
    pzt_els = np.split(nodes[n_plate:], 3)  # actuator, then the two sensors (4 hexahedra each)
    f = op.actuator_forces(pzt_els[0], coords, e, g, n, nz, case.V0, frame=1)
    W = np.array([op.sensor_weights(pe, coords, e, g, n, nz, frame=1) for pe in pzt_els[1:]])
    model = oso.SolidModel(nodes, coords, n, nz, prep["coef"], elem_mat=prep["elem_mat"])
    states = oa.forward_states(model, case.dt, prep["st"], f, 2, case.nft)
    ref = states[2:case.nft + 1] @ W.T  # voltage of u after the update at steps 2..nft
    assert np.abs(ref).max() > 0
    e_ = rel_l2(r["traces"][1:case.nft], ref)
    assert e_ < tol, e_
