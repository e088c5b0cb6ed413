"""GPU parity of the discrete adjoint (sem_adjoint) and of multi-signal excitation.

Bar: FP64 gradients within 1e-10 relative L2 of the oracle's adjoint (itself pinned by
finite differences on the CPU); gradients within 1e-4 of central finite differences of
the GPU's own forward traces; checkpoint segment length changes nothing (bitwise); the
split-precision form within 1e-3 of the FP64 oracle.
"""
import numpy as np
import pytest

import semhip
from cases import shell_pair, solid_pair, rel_l2
from oracle import adjoint as oa

pytestmark = pytest.mark.gpu


def _shell_setup(seed=0, quantity="vel", perturb=0.0):
    rng = np.random.default_rng(seed)
    dv = (4e3, 0, 4e3, 0, 2e3)
    mesh = dict(nex=4, ney=3, order=4, dv=dv, perturb=perturb, seed=7)  # one mesh for the three models
    base, inp = shell_pair(**mesh)
    ss = rng.uniform(0.8, 1.2, base.fen)
    rs = rng.uniform(0.8, 1.2, base.fen)
    model, _ = shell_pair(**mesh, stiff_scale=ss, rho_scale=rs)
    unit, _ = shell_pair(**mesh, stiff_scale=np.ones(base.fen), rho_scale=np.ones(base.fen))
    nsteps = 60
    st = rng.standard_normal(nsteps + 1)
    src = rng.choice(model.dof, 2, replace=False)
    amp = rng.uniform(0.5, 1.0, 2)
    ptr = np.array([0, 1, 4])
    dofs = np.concatenate([[5 * (model.n_nodes // 2) + 4], rng.choice(model.dof, 3, replace=False)])
    w = np.concatenate([[1.0], rng.uniform(0.5, 1.5, 3)])
    return dict(kind=semhip.SEM_SHELL5, model=model, unit=unit, inp=inp, ss=ss, rs=rs, dv=dv, st=st, src=src,
                amp=amp, ptr=ptr, dofs=dofs, w=w, nsteps=nsteps, dt=1e-7, quantity=quantity)


def _solid_setup(seed=0, quantity="disp_post", perturb=0.0):
    rng = np.random.default_rng(seed)
    dv = (3e3, 3e3, 1e3)
    mesh = dict(nex=3, ney=2, nlay=2, order=3, order_z=2, perturb=perturb, seed=7)  # one mesh for the three models
    base, inp = solid_pair(**mesh)
    ss = rng.uniform(0.8, 1.2, base.fen)
    rs = rng.uniform(0.8, 1.2, base.fen)
    model, _ = solid_pair(**mesh, dv=dv, stiff_scale=ss, rho_scale=rs)
    unit, _ = solid_pair(**mesh, dv=dv, stiff_scale=np.ones(base.fen), rho_scale=np.ones(base.fen))
    nsteps = 50
    st = rng.standard_normal(nsteps + 1)
    src = rng.choice(model.dof, 2, replace=False)
    amp = rng.uniform(0.5, 1.0, 2)
    ptr = np.array([0, 1, 3])
    dofs = np.concatenate([[3 * (model.n_nodes // 2) + 2], rng.choice(model.dof, 2, replace=False)])
    w = np.concatenate([[1.0], rng.uniform(0.5, 1.5, 2)])
    return dict(kind=semhip.SEM_SOLID3, model=model, unit=unit, inp=inp, ss=ss, rs=rs, dv=dv, st=st, src=src,
                amp=amp, ptr=ptr, dofs=dofs, w=w, nsteps=nsteps, dt=2e-8, quantity=quantity)


def _solver(c, ss=None, rs=None, kernel=(0, 0)):
    s = semhip.Solver(0)
    s.set_kernel(*kernel)
    inp = c["inp"]
    s.upload_mesh(c["kind"], inp["nodes"], inp["coords"], inp["n_xy"], inp.get("n_z", 1))
    s.set_material(inp["coef"], stiff_scale=c["ss"] if ss is None else ss, rho_scale=c["rs"] if rs is None else rs)
    s.set_damping(c["dv"])
    s.set_excitation(c["src"], c["amp"], c["st"])
    fun = [(c["dofs"][c["ptr"][j]:c["ptr"][j + 1]], c["w"][c["ptr"][j]:c["ptr"][j + 1]]) for j in range(len(c["ptr"]) - 1)]
    s.set_outputs(fun, c["quantity"])
    return s


def _oracle_grad(c, n0, dJ):
    f = np.zeros(c["model"].dof)
    f[c["src"]] = c["amp"]
    return oa.adjoint_gradients(c["model"], c["unit"], c["dt"], c["st"], f, n0, c["nsteps"], c["ptr"], c["dofs"],
                                c["w"], c["quantity"], dJ)


@pytest.mark.parametrize("kernel", [(0, 0), (1, 16)])
@pytest.mark.parametrize("quantity", ["disp", "vel", "disp_post", "acc"])
def test_shell_adjoint_matches_oracle(quantity, kernel):
    c = _shell_setup(seed=1, quantity=quantity)
    dJ = np.random.default_rng(11).standard_normal((c["nsteps"], 2))
    gs_o, gr_o = _oracle_grad(c, 1, dJ)
    with _solver(c, kernel=kernel) as s:
        g = s.adjoint(c["dt"], 1, c["nsteps"], dJ)
    assert np.abs(gs_o).max() > 0 and np.abs(gr_o).max() > 0
    assert rel_l2(g["g_stiff"], gs_o) < 1e-10, rel_l2(g["g_stiff"], gs_o)
    assert rel_l2(g["g_rho"], gr_o) < 1e-10, rel_l2(g["g_rho"], gr_o)


@pytest.mark.parametrize("kernel", [(0, 4), (1, 8), (3, 4), (3, 8)])
@pytest.mark.parametrize("quantity", ["disp_post", "acc"])
def test_solid_adjoint_matches_oracle(quantity, kernel):
    c = _solid_setup(seed=2, quantity=quantity)
    dJ = np.random.default_rng(12).standard_normal((c["nsteps"], 2))
    gs_o, gr_o = _oracle_grad(c, 2, dJ)
    with _solver(c, kernel=kernel) as s:
        g = s.adjoint(c["dt"], 2, c["nsteps"], dJ)
    assert rel_l2(g["g_stiff"], gs_o) < 1e-10, rel_l2(g["g_stiff"], gs_o)
    assert rel_l2(g["g_rho"], gr_o) < 1e-10, rel_l2(g["g_rho"], gr_o)


@pytest.mark.parametrize("kind,kernel", [("shell", (0, 16)), ("shell", (1, 8)), ("solid", (0, 4)), ("solid", (1, 8)), ("solid", (3, 4)), ("solid", (3, 8))])
def test_adjoint_on_distorted_mesh_matches_oracle(kind, kernel):
    """Non-affine elements: the forward and adjoint steps and the gradient epilogue read
    per-point geometry and quadrature weights."""
    if kind == "shell":
        c, n0 = _shell_setup(seed=5, quantity="vel", perturb=0.15), 1
    else:
        c, n0 = _solid_setup(seed=6, quantity="disp_post", perturb=0.15), 2
    dJ = np.random.default_rng(14).standard_normal((c["nsteps"], 2))
    gs_o, gr_o = _oracle_grad(c, n0, dJ)
    with _solver(c, kernel=kernel) as s:
        assert not s.is_affine()
        g = s.adjoint(c["dt"], n0, c["nsteps"], dJ)
    assert np.abs(gs_o).max() > 0 and np.abs(gr_o).max() > 0
    assert rel_l2(g["g_stiff"], gs_o) < 1e-10, rel_l2(g["g_stiff"], gs_o)
    assert rel_l2(g["g_rho"], gr_o) < 1e-10, rel_l2(g["g_rho"], gr_o)


def test_adjoint_segment_length_is_invisible_and_state_advances():
    c = _shell_setup(seed=3)
    dJ = np.random.default_rng(13).standard_normal((c["nsteps"], 2))
    out = []
    with _solver(c) as s:
        s.run(c["dt"], 1, c["nsteps"])
        u_run, uo_run = s.get_state()
        for seg in (0, 1, 7, c["nsteps"]):
            s.set_state(None, None)
            out.append(s.adjoint(c["dt"], 1, c["nsteps"], dJ, seg_len=seg))
            u, uo = s.get_state()
            assert np.array_equal(u, u_run) and np.array_equal(uo, uo_run)
    for o in out[1:]:
        assert np.array_equal(o["g_stiff"], out[0]["g_stiff"])
        assert np.array_equal(o["g_rho"], out[0]["g_rho"])


def test_adjoint_matches_gpu_finite_differences():
    c = _shell_setup(seed=4, quantity="disp")
    rng = np.random.default_rng(14)
    dJ = rng.standard_normal((c["nsteps"], 2))
    with _solver(c) as s:
        g = s.adjoint(c["dt"], 1, c["nsteps"], dJ)

        def J(ss, rs):
            s.set_material(c["inp"]["coef"], stiff_scale=ss, rho_scale=rs)
            s.set_state(None, None)
            return float((s.run(c["dt"], 1, c["nsteps"])["traces"] * dJ).sum())

        h = 1e-4
        for e in (0, 5, 11):
            for which, gv in (("s", g["g_stiff"]), ("r", g["g_rho"])):
                vals = []
                for sgn in (1, -1):
                    ss, rs = c["ss"].copy(), c["rs"].copy()
                    (ss if which == "s" else rs)[e] += sgn * h
                    vals.append(J(ss, rs))
                fd = (vals[0] - vals[1]) / (2 * h)
                assert abs(fd - gv[e]) <= 1e-4 * np.abs(gv).max(), (which, e, fd, gv[e])


def test_adjoint_fp32_split():
    c = _solid_setup(seed=5)
    dJ = np.random.default_rng(15).standard_normal((c["nsteps"], 2))
    gs_o, gr_o = _oracle_grad(c, 2, dJ)
    with _solver(c) as s:
        g = s.adjoint(c["dt"], 2, c["nsteps"], dJ, precision="fp32_split")
    assert rel_l2(g["g_stiff"], gs_o) < 1e-3, rel_l2(g["g_stiff"], gs_o)
    assert rel_l2(g["g_rho"], gr_o) < 1e-3, rel_l2(g["g_rho"], gr_o)


def test_adjoint_matrix_core_shell_kernel():
    """The gradient epilogue of the matrix-core shell kernel (FP32 form) against the FP64
    oracle's gradients."""
    c = _shell_setup(seed=8, quantity="disp")
    dJ = np.random.default_rng(18).standard_normal((c["nsteps"], 2))
    gs_o, gr_o = _oracle_grad(c, 1, dJ)
    with _solver(c, kernel=(2, 16)) as s:
        g = s.adjoint(c["dt"], 1, c["nsteps"], dJ, precision="fp32_split")
    assert rel_l2(g["g_stiff"], gs_o) < 1e-3, rel_l2(g["g_stiff"], gs_o)
    assert rel_l2(g["g_rho"], gr_o) < 1e-3, rel_l2(g["g_rho"], gr_o)


def test_torch_autograd_wrapper():
    import torch
    from semhip.autograd import TraceSimulation
    c = _shell_setup(seed=6, quantity="vel")
    with _solver(c) as s:
        sim = TraceSimulation(s, c["inp"]["coef"], c["dt"], 1, c["nsteps"])
        ss = torch.tensor(c["ss"], dtype=torch.float64, requires_grad=True)
        rs = torch.tensor(c["rs"], dtype=torch.float64, requires_grad=True)
        tr = sim.traces(ss, rs)
        assert tr.shape == (c["nsteps"], 2)
        obs = tr.detach() * 0.9
        loss = 0.5 * ((tr - obs) ** 2).sum()
        loss.backward()
        gs_o, gr_o = _oracle_grad(c, 1, (tr - obs).detach().numpy())
    assert rel_l2(ss.grad.numpy(), gs_o) < 1e-10
    assert rel_l2(rs.grad.numpy(), gr_o) < 1e-10


def test_multi_signal_excitation_superposes():
    """Two sources with their own signals equal the sum of two single-signal runs of the
    oracle (the loop is linear and starts at rest)."""
    c = _shell_setup(seed=7, quantity="disp")
    rng = np.random.default_rng(17)
    st2 = np.stack([c["st"], rng.standard_normal(c["st"].size)])
    outs = list(c["dofs"][:3])
    ref = 0
    for k in range(2):
        f = np.zeros(c["model"].dof)
        f[c["src"][k]] = c["amp"][k]
        ref = ref + c["model"].run(c["dt"], st2[k], f, 1, c["nsteps"], outs, "disp")["traces"][:c["nsteps"]]
    with _solver(c) as s:
        s.set_excitation(c["src"], c["amp"], st2, sig=[0, 1])
        s.set_outputs(outs, "disp")
        r = s.run(c["dt"], 1, c["nsteps"])
    assert np.abs(ref).max() > 0
    assert rel_l2(r["traces"], ref) < 1e-10
