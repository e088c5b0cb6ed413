"""GPU parity of the 3D-solid path against the oracle (FP64 1e-10, FP32 split 1e-3)."""
import numpy as np
import pytest

import semhip
from cases import solid_pair, rel_l2
from oracle.common import hanning_signal

pytestmark = pytest.mark.gpu

# element kernels: (variant, elements per patch)
KERNELS = [(0, 4), (1, 4), (1, 8), (3, 4), (3, 8)]


def _solver(inp, kernel=(0, 4)):
    s = semhip.Solver(0)
    s.set_kernel(*kernel)
    s.upload_mesh(semhip.SEM_SOLID3, inp["nodes"], inp["coords"], inp["n_xy"], inp["n_z"])
    s.set_material(inp["coef"])
    assert s.kernel() == kernel
    return s


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("order,order_z", [(2, 2), (3, 3), (4, 4), (5, 5), (5, 2), (3, 2)])
def test_apply_K_orders(order, order_z, kernel):
    model, inp = solid_pair(nex=2, ney=2, nlay=2, order=order, order_z=order_z)
    with _solver(inp, kernel) as s:
        assert s.is_affine()
        u = np.random.default_rng(order).standard_normal(model.dof)
        # 8-element order-5 patches in FP64 use ~128 KB of LDS per workgroup (one per CU)
        assert rel_l2(s.apply_K(u), model.internal_force(u)) < 1e-12
        assert rel_l2(s.apply_K(u, "fp32_split"), model.internal_force(u)) < 1e-4


@pytest.mark.parametrize("kernel", KERNELS)
def test_apply_K_distorted(kernel):
    model, inp = solid_pair(nex=3, ney=2, nlay=2, order=4, perturb=0.2, seed=5)
    with _solver(inp, kernel) as s:
        assert not s.is_affine()
        u = np.random.default_rng(2).standard_normal(model.dof)
        ref = model.internal_force(u)
        assert rel_l2(s.apply_K(u), ref) < 1e-12
        # the FP32 form reads the per-point geometry as float
        assert rel_l2(s.apply_K(u, "fp32_split"), ref) < 1e-4


@pytest.mark.parametrize("kernel", KERNELS)
def test_apply_K_fp32(kernel):
    model, inp = solid_pair(nex=2, ney=2, nlay=2, order=4)
    with _solver(inp, kernel) as s:
        u = np.random.default_rng(3).standard_normal(model.dof)
        assert rel_l2(s.apply_K(u, "fp32_split"), model.internal_force(u)) < 1e-4


def test_kernel_variants_agree_on_a_run():
    """Every element kernel gives the same traces to rounding on a damped, excited run."""
    model, inp = solid_pair(nex=4, ney=3, nlay=2, order=4, lx=0.02, ly=0.015)
    dt, nft = 2e-8, 120
    _, st = hanning_signal(dt, nft, 100e3 / 5, 100e3, 0.0)
    out = []
    for kernel in KERNELS:
        with _solver(inp, kernel) as s:
            s.set_damping([1e4, 1e4, 1e4])
            s.set_excitation([3 * (model.n_nodes // 2) + 2], [1.0], st)
            s.set_outputs([3 * (model.n_nodes // 3) + 2, 3 * (model.n_nodes // 2) + 2], "disp_post")
            out.append(s.run(dt, 2, nft)["traces"])
    for o in out[1:]:
        assert rel_l2(o, out[0]) < 1e-12


def test_mass_matches_oracle():
    model, inp = solid_pair(nex=2, ney=3, nlay=2, order=4)
    with _solver(inp) as s:
        assert rel_l2(s.mass(), model.mass) < 1e-13
        IG, IL = s.baskets()
        assert np.array_equal(IG, model.I_G + 1)
        assert np.array_equal(IL, model.I_L + 1)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_run_matches_oracle(precision, tol, kernel):
    model, inp = solid_pair(nex=4, ney=4, nlay=2, order=4, lx=0.02, ly=0.02)
    dt, nft = 2e-8, 400
    _, st = hanning_signal(dt, nft, 100e3 / 5, 100e3, 0.0)
    top = np.nonzero(np.isclose(inp["coords"][:, 2], inp["coords"][:, 2].max()))[0]
    d = inp["coords"][top, :2] - 0.01
    src = 3 * top[np.argmin((d ** 2).sum(1))] + 2
    outs = [3 * top[k] + 2 for k in (0, len(top) // 3, len(top) // 2)]
    f_ext = np.zeros(model.dof)
    f_ext[src] = 1.0
    o = model.run(dt, st, f_ext, 2, nft, outs, "disp_post", [100, 300], "disp_post")
    with _solver(inp, kernel) as s:
        s.set_excitation([src], [1.0], st)
        s.set_outputs(outs, "disp_post")
        s.set_frames([100, 300], "disp_post")
        r = s.run(dt, 2, nft, precision)
    assert np.abs(o["traces"]).max() > 0
    e = rel_l2(r["traces"], o["traces"])
    assert e < tol, e
    assert rel_l2(r["frames"], o["frames"]) < tol


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_run_distorted_matches_oracle(precision, tol, kernel):
    """A damped run on a distorted (non-affine) mesh: per-point geometry in both forms."""
    model, inp = solid_pair(nex=4, ney=3, nlay=2, order=4, lx=0.02, ly=0.015, perturb=0.15, seed=11,
                            dv=(1e4, 1e4, 1e4))
    dt, nft = 2e-8, 300
    _, st = hanning_signal(dt, nft, 100e3 / 5, 100e3, 0.0)
    src = 3 * (model.n_nodes // 2) + 2
    outs = [3 * (model.n_nodes // 3) + 2, 3 * (model.n_nodes // 4) + 1, src]
    f_ext = np.zeros(model.dof)
    f_ext[src] = 1.0
    o = model.run(dt, st, f_ext, 2, nft, outs, "vel", [], "vel")
    with _solver(inp, kernel) as s:
        assert not s.is_affine()
        s.set_damping([1e4, 1e4, 1e4])
        s.set_excitation([src], [1.0], st)
        s.set_outputs(outs, "vel")
        r = s.run(dt, 2, nft, precision)
    assert np.abs(o["traces"]).max() > 0
    e = rel_l2(r["traces"], o["traces"])
    assert e < tol, e


@pytest.mark.parametrize("kernel", KERNELS)
def test_layered_laminate_matches_oracle(kernel):
    """One element layer per ply, each with its own rotated stiffness (composite3D_transf.m:109-119,
    prop='layered'; the engine's SolidCase(prop="layered")): the library's per-ply rows equal the
    oracle's, and the operator (FP64 1e-12, FP32 form 1e-4), the lumped mass and a point-force run
    (FP64 1e-10, FP32 form 1e-3) equal the oracle's with the same rows and element layers."""
    from semhip import engine as E
    from oracle import adjoint as oa
    from oracle import solid as oso
    case = E.SolidCase(lx=0.03, ly=0.02, nex=3, ney=2, nlay_el=4, h=2e-3, order=3, order_z=3, lh=(0.5e-3,) * 4,
                       angles=(0, 45, -45, 90), prop="layered", nft=150, force_xy=(0.015, 0.01),
                       sensors=((0.005, 0.005), (0.025, 0.015)))
    s, prep = E.solid_solver(case, kernel=kernel)
    n, nz = case.order + 1, case.order_z + 1
    Q = oso.lamina_q3d(case.E11, case.E22, case.E22, case.G12, case.G13, case.G23, case.nu12, case.nu12,
                       case.E22 / (2 * case.G23) - 1.0)
    rows = oso.laminate_solid_layered(case.angles, case.rho, **Q)
    coef = np.array([[r[k] for k in oso.SOLID_COEFS] for r in rows])
    assert np.abs(prep["coef"] - coef).max() <= 1e-12 * np.abs(coef).max()
    assert np.array_equal(prep["elem_mat"], np.repeat(np.arange(4), case.nex * case.ney))
    model = oso.SolidModel(prep["nodes"], prep["coords"], n, nz, coef, elem_mat=prep["elem_mat"])
    u = np.random.default_rng(7).standard_normal(model.dof)
    ref = model.internal_force(u)
    assert rel_l2(s.apply_K(u), ref) < 1e-12
    assert rel_l2(s.apply_K(u, "fp32_split"), ref) < 1e-4
    assert rel_l2(s.mass(), model.mass) < 1e-13
    f = np.zeros(model.dof)
    f[prep["ex_dofs"]] = prep["ex_amp"]
    states = oa.forward_states(model, case.dt, prep["st"], f, 2, case.nft)
    tr_ref = states[2:case.nft + 1][:, np.asarray(prep["outputs"])]
    assert np.abs(tr_ref).max() > 0
    for precision, tol in (("fp64", 1e-10), ("fp32_split", 1e-3)):
        s.set_state(None, None)
        r = s.run(case.dt, 2, case.nft, precision)
        assert rel_l2(r["traces"][1:case.nft], tr_ref) < tol, precision
    s.close()
