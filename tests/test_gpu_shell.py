"""GPU parity of the flat-shell path (libsemhip.so through its C-ABI) against the oracle.

Bar: FP64 operator and traces within 1e-10 relative L2 of the oracle (the path's
stated FP64 tolerance); integer maps bit-exact; reruns bitwise identical; the
split-precision form within 1e-3 relative L2 of the FP64 oracle.
"""
import numpy as np
import pytest

import semhip
from semhip import engine as E
from cases import shell_pair, rel_l2
from oracle import common as oc

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_TOL = 1e-3


# element kernels under test: the default (variant 0) and the line-wise kernel (variant 1)
# with 16- and 8-element patches
LINE = [(1, 16), (1, 8)]


def _solver(inp, model, kernel=None):
    s = semhip.Solver(0)
    if kernel is not None:
        s.set_kernel(*kernel)
    s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], inp["n_xy"])
    s.set_material(inp["coef"])
    return s


@pytest.mark.parametrize("kernel", [None] + LINE)
@pytest.mark.parametrize("order", [2, 3, 4, 5, 6, 7])
def test_apply_K_affine_orders(order, kernel):
    model, inp = shell_pair(nex=3, ney=2, order=order)
    with _solver(inp, model, kernel) as s:
        assert s.is_affine()
        rng = np.random.default_rng(order)
        for _ in range(3):
            u = rng.standard_normal(model.dof)
            f = s.apply_K(u)
            ref = model.internal_force(u)
            assert rel_l2(f, ref) < 1e-12, order


@pytest.mark.parametrize("kernel", [None] + LINE)
def test_apply_K_distorted_mesh(kernel):
    model, inp = shell_pair(nex=3, ney=3, order=5, perturb=0.2, seed=3)
    with _solver(inp, model, kernel) as s:
        assert not s.is_affine()
        u = np.random.default_rng(1).standard_normal(model.dof)
        ref = model.internal_force(u)
        assert rel_l2(s.apply_K(u), ref) < 1e-12
        # the FP32 form reads the per-point geometry as float
        assert rel_l2(s.apply_K(u, "fp32_split"), ref) < 1e-4


@pytest.mark.parametrize("kernel", [None] + LINE)
def test_apply_K_fp32_split(kernel):
    model, inp = shell_pair(nex=3, ney=2, order=5)
    with _solver(inp, model, kernel) as s:
        u = np.random.default_rng(7).standard_normal(model.dof)
        e = rel_l2(s.apply_K(u, "fp32_split"), model.internal_force(u))
        assert e < 1e-4


def test_mass_and_baskets_match_oracle():
    model, inp = shell_pair(nex=4, ney=3, order=5)
    with _solver(inp, model) as s:
        assert rel_l2(s.mass(), model.mass) < 1e-13
        IG, IL = s.baskets()
        assert np.array_equal(IG, model.I_G + 1)
        assert np.array_equal(IL, model.I_L + 1)


def _run_pair(case, nsteps, precision="fp64", kernel=None):
    prep = E.prepare_flat_shell(case)
    s, _ = E.flat_shell_solver(case, prep, kernel=kernel)
    r = s.run(case.dt, 1, nsteps, precision)
    s.close()
    return prep, r


def _oracle_run(case, prep, nsteps):
    from oracle import shell as osh
    Q = osh.lamina_q(case.E11, case.E22, case.G12, case.G13, case.G23, case.nu12)
    lam = osh.laminate_shell(list(case.lh), list(case.angles), case.rho, **Q)
    coef = np.array([[lam[k] for k in osh.SHELL_COEFS]])
    model = osh.ShellModel(prep["nodes"], prep["coords"], case.order, coef, stiff_scale=case.stiff_scale,
                           rho_scale=case.rho_scale,
                           dv_field=(case.etad_xy, 0, case.etad_xy, 0, case.etad_z))
    f_ext = np.zeros(model.dof)
    f_ext[prep["force_dof"]] = case.V0
    out = model.run(case.dt, prep["st"], f_ext, 1, nsteps, prep["sensor_dofs"], case.out_quantity,
                    [f for f in prep["frame_no"] if f <= nsteps], case.field_variable,
                    h_guard=prep["h"])
    return out


def small_case(**kw):
    base = dict(lx=0.1, ly=0.1, nex=8, ney=8, order=5, force_xy=(0.05, 0.05),
                sensors=((0.08, 0.05), (0.05, 0.08), (0.02, 0.05), (0.05, 0.02)), nft=600,
                n_frames=4, dt=1e-7)
    base.update(kw)
    return E.shell_plate_40x40(**base)


@pytest.mark.parametrize("kernel", [None, (1, 16)])
def test_run_traces_and_frames_match_oracle(kernel):
    case = small_case()
    prep, r = _run_pair(case, case.nft, kernel=kernel)
    o = _oracle_run(case, prep, case.nft)
    assert r["status"] == "ok" and r["nn_done"] == case.nft
    assert np.abs(o["traces"]).max() > 0
    assert rel_l2(r["traces"], o["traces"]) < FP64_TOL
    assert rel_l2(r["frames"], o["frames"]) < FP64_TOL
    assert np.allclose(r["t_frames"], o["t_frames"], rtol=0, atol=1e-18)


@pytest.mark.parametrize("kernel", [None, (1, 16)])
def test_run_displacement_outputs_and_damping(kernel):
    case = small_case(out_quantity="disp", field_variable="displacement", etad_xy=2e3, etad_z=5e3,
                      nft=300)
    prep, r = _run_pair(case, case.nft, kernel=kernel)
    o = _oracle_run(case, prep, case.nft)
    assert rel_l2(r["traces"], o["traces"]) < FP64_TOL
    assert rel_l2(r["frames"], o["frames"]) < FP64_TOL


@pytest.mark.parametrize("kernel", [None, (1, 8)])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32_split", FP32_TOL)])
def test_acceleration_frames_match_oracle(kernel, precision, tol):
    """Frames of the acceleration (u_old - 2 U + u_new) / dt^2 (force_excit_c.m:710-730)."""
    case = small_case(field_variable="acceleration", nft=400)
    prep, r = _run_pair(case, case.nft, precision, kernel=kernel)
    o = _oracle_run(case, prep, case.nft)
    assert np.abs(o["frames"]).max() > 0
    assert rel_l2(r["frames"], o["frames"]) < tol
    assert rel_l2(r["traces"], o["traces"]) < tol


@pytest.mark.parametrize("kernel", [None, (1, 8)])
def test_run_with_stiffness_and_density_fields(kernel):
    rng = np.random.default_rng(11)
    ss = rng.uniform(0.5, 1.0, 64)
    rs = rng.uniform(0.8, 1.2, 64)
    case = small_case(stiff_scale=ss, rho_scale=rs, nft=300)
    prep, r = _run_pair(case, case.nft, kernel=kernel)
    o = _oracle_run(case, prep, case.nft)
    assert rel_l2(r["traces"], o["traces"]) < FP64_TOL


@pytest.mark.parametrize("kernel", [None, (1, 16)])
def test_rerun_bitwise_identical(kernel):
    case = small_case(nft=200)
    _, r1 = _run_pair(case, case.nft, kernel=kernel)
    _, r2 = _run_pair(case, case.nft, kernel=kernel)
    assert np.array_equal(r1["traces"], r2["traces"])
    assert np.array_equal(r1["frames"], r2["frames"])


def test_split_run_resume_equals_single_run():
    case = small_case(nft=240)
    prep = E.prepare_flat_shell(case)
    s, _ = E.flat_shell_solver(case, prep)
    ra = s.run(case.dt, 1, 240)
    s.close()
    s, _ = E.flat_shell_solver(case, prep)
    rb1 = s.run(case.dt, 1, 97, ld=240)
    rb2 = s.run(case.dt, 98, 240, ld=240)
    s.close()
    t = rb1["traces"].copy()
    t[97:] = rb2["traces"][97:]
    assert np.array_equal(t, ra["traces"])


@pytest.mark.parametrize("kernel", [(0, 16), (2, 16)] + LINE)
def test_fp32_split_run_error(kernel):
    case = small_case()
    prep, r = _run_pair(case, case.nft, "fp32_split", kernel)
    o = _oracle_run(case, prep, case.nft)
    e = rel_l2(r["traces"], o["traces"])
    assert e < FP32_TOL, e
    assert rel_l2(r["frames"], o["frames"]) < FP32_TOL


@pytest.mark.parametrize("order", [4, 5])
def test_matrix_core_kernel_operator(order):
    """The bf16 hi/lo matrix-core contractions: operator within 1e-4 of the FP64 oracle on
    regular and distorted meshes, with a stiffness field; FP64 is refused for this kernel."""
    ss = np.random.default_rng(1).uniform(0.5, 1.0, 20)
    for perturb in (0.0, 0.2):
        model, inp = shell_pair(nex=5, ney=4, order=order, perturb=perturb, seed=order, stiff_scale=ss)
        with semhip.Solver(0) as s:
            s.set_kernel(2, 16)
            s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], inp["n_xy"])
            s.set_material(inp["coef"], stiff_scale=ss)
            u = np.random.default_rng(order).standard_normal(model.dof)
            e = rel_l2(s.apply_K(u, "fp32_split"), model.internal_force(u))
            assert e < 1e-4, (perturb, e)
            with pytest.raises(semhip.SemError):
                s.apply_K(u, "fp64")


@pytest.mark.parametrize("kernel,precision,tol", [((0, 16), "fp64", FP64_TOL), ((1, 16), "fp64", FP64_TOL),
                                                   ((1, 8), "fp64", FP64_TOL), ((0, 16), "fp32_split", FP32_TOL),
                                                   ((1, 8), "fp32_split", FP32_TOL),
                                                   ((2, 16), "fp32_split", FP32_TOL)])
def test_run_distorted_mesh_matches_oracle(kernel, precision, tol):
    """A damped run on a distorted (non-affine) mesh, per-point geometry in both forms."""
    dv = (2e3, 0, 2e3, 0, 5e3)
    model, inp = shell_pair(nex=4, ney=3, order=5, perturb=0.15, seed=2, dv=dv)
    dt, nft = 1e-7, 400
    _, st = oc.hanning_signal(dt, nft, 50e3 / 5, 50e3, 0.0)
    src = 5 * (model.n_nodes // 2) + 4  # transverse force, transverse velocity sensors
    outs = [5 * (model.n_nodes // 3) + 4, 5 * (2 * model.n_nodes // 3) + 4, src]
    f_ext = np.zeros(model.dof)
    f_ext[src] = 1.0
    o = model.run(dt, st, f_ext, 1, nft, outs, "vel", [], "vel")
    with _solver(inp, model, kernel) as s:
        assert not s.is_affine()
        s.set_damping(dv)
        s.set_excitation([src], [1.0], st)
        s.set_outputs(outs, "vel")
        r = s.run(dt, 1, nft, precision)
    assert np.abs(o["traces"]).max() > 0
    e = rel_l2(r["traces"], o["traces"])
    assert e < tol, e


def test_divergence_guard_matches_oracle():
    # dt far above the CFL limit: the guard trips at the first frame step, as in the reference
    case = small_case(dt=2e-6, nft=400, n_frames=8)
    prep, r = _run_pair(case, case.nft)
    o = _oracle_run(case, prep, case.nft)
    assert o["status"] == "diverged"
    assert r["status"] == "diverged"
    assert r["nn_done"] == o["nn_done"]
    n = o["nn_done"]
    assert np.all(r["traces"][n:] == 0)
    fin = np.isfinite(o["traces"][:n]).all(axis=1)
    assert rel_l2(r["traces"][:n][fin], o["traces"][:n][fin]) < 1e-8


@pytest.mark.parametrize("kernel", [None, (1, 16)])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("fp32_split", FP32_TOL)])
def test_delaminated_plate_matches_oracle(precision, tol, kernel):
    """A delaminated plate (node split after ply 3, sub-laminates above and under) against
    the oracle on the same split mesh and coefficient table."""
    from oracle import shell as osh
    case = small_case(delam_rect=(0.03, 0.03, 0.07, 0.06), delamination_layer=3, nft=400)
    prep = E.prepare_flat_shell(case)
    assert prep["coords"].shape[0] > (8 * 5 + 1) ** 2 and len(prep["den_above"]) == 12
    s, _ = E.flat_shell_solver(case, prep, kernel=kernel)
    r = s.run(case.dt, 1, case.nft, precision)
    s.close()
    model = osh.ShellModel(prep["nodes"], prep["coords"], case.order, prep["coef"], elem_mat=prep["elem_mat"])
    f_ext = np.zeros(model.dof)
    f_ext[prep["force_dof"]] = case.V0
    o = model.run(case.dt, prep["st"], f_ext, 1, case.nft, prep["sensor_dofs"], case.out_quantity,
                  [f for f in prep["frame_no"] if f <= case.nft], case.field_variable, h_guard=prep["h"])
    assert r["status"] == "ok" and np.abs(o["traces"]).max() > 0
    assert rel_l2(r["traces"], o["traces"]) < tol
    # the delamination changes the response (against the pristine plate)
    p0 = E.prepare_flat_shell(small_case(nft=400))
    s0, _ = E.flat_shell_solver(small_case(nft=400), p0)
    r0 = s0.run(case.dt, 1, case.nft, "fp64")
    s0.close()
    assert rel_l2(r0["traces"], r["traces"]) > 1e-3


def test_engine_writes_reference_output_files(tmp_path):
    """main_flat_shell_force_excit_c with an output folder writes displ<k>.mat,
    time<k>.mat and the frame files with the run's own arrays."""
    import scipy.io as sio
    case = small_case(nft=200, n_frames=2)
    out = E.main_flat_shell_force_excit_c(case, output_name=tmp_path, test_case=7)
    assert out["status"] == "ok"
    assert np.array_equal(sio.loadmat(str(tmp_path / "displ7.mat"))["displ"], out["displ"])
    tm = sio.loadmat(str(tmp_path / "time7.mat"))
    assert tm["TotalNofNodes"].item() == (8 * 5 + 1) ** 2
    for c in (1, 2):
        vz = sio.loadmat(str(tmp_path / f"Vz_frame{c}.mat"))["Vz"].reshape(-1)
        assert np.array_equal(vz, out["frames"][c - 1][4::5])


@pytest.mark.parametrize("kernel,precision,tol", [(None, "fp64", FP64_TOL), ((1, 8), "fp64", FP64_TOL),
                                                   (None, "fp32_split", FP32_TOL), ((1, 8), "fp32_split", FP32_TOL),
                                                   ((2, 16), "fp32_split", FP32_TOL)])
def test_gmsh_fixture_mesh_run_matches_oracle(kernel, precision, tol):
    """A run on the reference-held gmsh mesh (gmsh_out/m1_rand_2.m: 16 unstructured quads,
    some given clockwise) meshed by the library (change_turn_quad, quad2spectral, order 5):
    distorted elements, per-point geometry, against the oracle on the oracle's own mesh of the
    same quads.  The step is half the critical step of the mesh (power iteration on M^-1 K)."""
    import json
    from pathlib import Path
    from oracle import mesh as om, shell as osh
    m = json.loads((Path(__file__).resolve().parent / "golden" / "m1_rand_2.json").read_text())
    pos = np.array(m["POS"])
    quads = np.array(m["QUADS"])[:, :4] - 1
    quads[1::4] = quads[1::4][:, [0, 3, 2, 1]]
    q, _ = semhip.change_turn_quad(quads, pos)
    nodes, coords = semhip.quad2spectral(q, pos, 5)
    n_o, c_o = om.quad2spectral(om.change_turn_quad(quads, pos), pos, 5)
    Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    lam = osh.laminate_shell([0.25e-3] * 8, [0, 90, 0, 90, 90, 0, 90, 0], 1570, **Q)
    coef = np.array([[lam[k] for k in osh.SHELL_COEFS]])
    model = osh.ShellModel(n_o, c_o, 5, coef)
    x = np.random.default_rng(0).standard_normal(model.dof)
    for _ in range(100):
        y = model.internal_force(x) / model.mass
        lmax = np.linalg.norm(y) / np.linalg.norm(x)
        x = y / np.linalg.norm(y)
    dt, nft = 0.5 * 2 / np.sqrt(lmax), 400
    _, st = oc.hanning_signal(dt, nft, 500e3 / 5, 500e3, 0.0)
    src = 5 * int(np.argmin(np.hypot(coords[:, 0] - 0.25, coords[:, 1] - 0.25))) + 4
    outs = [5 * int(np.argmin(np.hypot(coords[:, 0] - x0, coords[:, 1] - y0))) + 4
            for x0, y0 in ((0.253, 0.25), (0.247, 0.252))]
    s = semhip.Solver(0)
    if kernel is not None:
        s.set_kernel(*kernel)
    s.upload_mesh(semhip.SEM_SHELL5, nodes, coords, 6)
    assert not s.is_affine()
    s.set_material(coef)
    s.set_excitation([src], [1.0], st)
    s.set_outputs(outs, "vel")
    r = s.run(dt, 1, nft, precision)
    s.close()
    f = np.zeros(model.dof)
    f[src] = 1.0
    o = model.run(dt, st, f, 1, nft, outs, "vel")
    assert np.abs(o["traces"]).max() > 0
    e = rel_l2(r["traces"][:nft], o["traces"][:nft])
    assert e < tol, e


def fan_quads(n_blades, r=0.01, rm=0.013):
    """A disk of n_blades kite-shaped quads around one centre node (counter-clockwise):
    blade k is (centre, v_k, m_k, v_k+1) with v on a circle of radius r and m, between
    them, on radius rm > r cos(pi / n_blades), so every blade is convex."""
    ang = 2 * np.pi * np.arange(n_blades) / n_blades
    v = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(n_blades)], axis=1)
    m = np.stack([rm * np.cos(ang + np.pi / n_blades), rm * np.sin(ang + np.pi / n_blades), np.zeros(n_blades)],
                 axis=1)
    pos = np.concatenate([np.zeros((1, 3)), v, m])
    k = np.arange(n_blades)
    quads = np.stack([np.zeros(n_blades, int), 1 + k, 1 + n_blades + k, 1 + (k + 1) % n_blades], axis=1)
    return quads.astype(np.int64), pos


@pytest.mark.parametrize("kernel", [None, (1, 16), (1, 8)])
@pytest.mark.parametrize("order", [2, 5])
def test_apply_K_node_shared_by_ten_elements(order, kernel):
    """The centre of a ten-blade fan occurs in ten elements of one 16-element patch, more
    than a patch node's occurrence record holds (8): the assembly takes the first seven
    from the record and the rest from the node's occurrence list, in list order.  With
    8-element patches the centre is a boundary node of two patches instead."""
    from oracle import mesh as om
    from oracle import shell as osh
    q, p = fan_quads(10)
    nodes, coords = om.quad2spectral(q, p, order)
    Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    lam = osh.laminate_shell([0.25e-3] * 8, [0, 90, 0, 90, 90, 0, 90, 0], 1570, **Q)
    coef = np.array([[lam[k] for k in osh.SHELL_COEFS]])
    model = osh.ShellModel(nodes, coords, order, coef)
    inp = dict(nodes=nodes, coords=coords, coef=coef, n_xy=order + 1)
    with _solver(inp, model, kernel) as s:
        rng = np.random.default_rng(order)
        for _ in range(2):
            u = rng.standard_normal(model.dof)
            ref = model.internal_force(u)
            assert rel_l2(s.apply_K(u), ref) < 1e-12
            assert rel_l2(s.apply_K(u, "fp32_split"), ref) < 1e-4
