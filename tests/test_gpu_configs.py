"""The five benchmark workloads on the GPU at their own sizes and orders.

* 40 x 40 order-5 flat shell (the whole plate): 2000 steps (the wave has reached the four
  sensors 0.15 m away) and the four snapshots in them against the oracle (FP64 1e-10, FP32
  form 1e-3);
* the 3D PZT plates at their real orders and laminates -- order 4 with PZT voltage sensors
  ([0/90/0/90]s, 100 kHz) and order 5 with surface w sensors and snapshots ([0/45/-45/90]2s,
  75 kHz) -- on reduced plates against the oracle (both also over 4000 steps, the whole
  burst and its reflections, in both forms), and at full size (100 x 100 x 2 and
  250 x 250 x 2 elements) through properties the oracle cannot reach: a rerun is bitwise
  the same, twice the drive gives twice the traces, the FP32 form stays within 1e-3;
* the damage sweep: cases of 60, 90 and 120 elements per side stepped as one batch equal
  their own single-case runs, the 60 x 60 case equals the oracle, the FP32 batch stays
  within 1e-3 per case;
* the 120 x 120 adjoint inversion over its 15000 steps: the adjoint gradient equals central
  differences of the misfit (FP64 forward runs) on the elements with the largest gradients
  and random ones.
"""
import time

import numpy as np
import pytest

from semhip import engine as E
from semhip import batch as B
from semhip import inversion as I
from semhip import pzt
from cases import rel_l2
from oracle import adjoint as oa
from oracle import pzt as op
from oracle import shell as osh
from oracle import solid as oso

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


_ORACLE = {}  # oracle results shared by the parametrisations of one workload


def cached(key, fn):
    if key not in _ORACLE:
        _ORACLE[key] = fn()
    return _ORACLE[key]


# ---------------------------------------------------------------------------- 40 x 40 shell
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_shell_40x40_whole_plate_matches_oracle(precision, tol):
    case = E.shell_plate_40x40()  # 5000-step schedule: snapshots every 500 steps
    prep = E.prepare_flat_shell(case)
    assert prep["frame_no"][:4] == [500, 1000, 1500, 2000] and prep["coords"].shape[0] * 5 == 202005
    n = 2000
    s, _ = E.flat_shell_solver(case, prep)
    r = s.run(case.dt, 1, n, precision)
    s.close()
    model = osh.ShellModel(prep["nodes"], prep["coords"], case.order, np.atleast_2d(prep["coef"]))
    f = np.zeros(model.dof)
    f[prep["force_dof"]] = case.V0
    t0 = time.perf_counter()
    o = cached("shell40", lambda: model.run(case.dt, prep["st"], f, 1, n, prep["sensor_dofs"], "vel",
                                            prep["frame_no"][:4], "velocity"))
    log(f"oracle 40x40, {n} steps: {time.perf_counter() - t0:.1f} s")
    ref = o["traces"][:n]
    assert np.abs(ref).max() > 0 and np.abs(o["frames"][0]).max() > 0
    e_tr = rel_l2(r["traces"][:n], ref)
    e_fr = max(rel_l2(r["frames"][i], o["frames"][i]) for i in range(4))
    log(f"40x40 {precision}: traces {e_tr:.2e}, snapshot {e_fr:.2e}")
    assert e_tr < tol and e_fr < tol


# ---------------------------------------------------------------------------- 3D PZT plates
def _condensed(v, col, n_nodes):
    out = np.zeros(3 * n_nodes)
    np.add.at(out.reshape(-1, 3), col, v.reshape(-1, 3))
    return out


def _solid_oracle_model(case, prep):
    """The oracle's model of a 3D PZT plate and its actuator forces: the oracle's actuator on
    the same surface transducer mesh, carried down its columns."""
    nodes, coords = prep["nodes"], prep["coords"]
    n, nz = case.order + 1, case.order_z + 1
    _, e, g, _ = pzt.pzt5a()
    pn, X, col = pzt.surface_transducer(nodes, coords, n, nz, _footprint(case, prep, case.force_xy), case.pzt_t)
    f = _condensed(op.actuator_forces(pn, X, e, g, n, nz, case.V0, frame=1), col, coords.shape[0])
    return oso.SolidModel(nodes, coords, n, nz, np.atleast_2d(prep["coef"])), f


def _footprint(case, prep, xy):
    hx = case.lx / case.nex
    c = (round(xy[0] / hx) * hx, round(xy[1] / hx) * hx)
    return pzt.top_elements(prep["nodes"], prep["coords"], c, case.pzt_size / 2, case.order + 1, case.order_z + 1)


def _solid_oracle(case, prep, nn_last):
    """Oracle states of a 3D PZT plate; returns (states u_1..u_{nn_last+1}, sensor weights or
    None for point sensors)."""
    model, f = _solid_oracle_model(case, prep)
    t0 = time.perf_counter()
    states = oa.forward_states(model, case.dt, prep["st"], f, 2, nn_last)
    log(f"oracle {case.nex}x{case.ney}x{case.nlay_el} order {case.order}: {time.perf_counter() - t0:.1f} s")
    return states, (_voltage_weights(case, prep) if case.sensor_kind == "pzt" else None)


def _voltage_weights(case, prep):
    """The oracle's PZT sensors on the same surface transducer meshes: their voltage as
    weights over the DOFs [n_sensors, dof]."""
    nodes, coords = prep["nodes"], prep["coords"]
    n, nz = case.order + 1, case.order_z + 1
    _, e, g, _ = pzt.pzt5a()
    W = []
    for xy in case.sensors:
        pn_s, X_s, col_s = pzt.surface_transducer(nodes, coords, n, nz, _footprint(case, prep, xy), case.pzt_t)
        W.append(_condensed(op.sensor_weights(pn_s, X_s, e, g, n, nz, frame=1), col_s, coords.shape[0]))
    return np.array(W)


@pytest.mark.parametrize("kernel", [(1, 4), (1, 8), (3, 4), (3, 8)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_solid_order4_pzt_voltage_plate_matches_oracle(kernel, precision, tol):
    """The 100 x 100 x 2 workload's element (order 4, [0/90/0/90]s, 2 layers), actuator
    (10 mm PZT-5A, 100 kHz, 10 V) and PZT voltage sensors on a 12 x 12 x 2 plate."""
    case = E.solid_plate_100x100_pzt(lx=0.06, ly=0.06, nex=12, ney=12, nft=250, force_xy=(0.03, 0.03),
                                     sensors=((0.045, 0.03), (0.03, 0.045)))
    s, prep = E.solid_solver(case, kernel=kernel)
    r = s.run(case.dt, 2, case.nft, precision)
    s.close()
    states, W = cached("solid_o4", lambda: _solid_oracle(case, prep, case.nft))
    ref = states[2:case.nft + 1] @ W.T  # voltage of u after the update at steps 2..nft
    assert np.abs(ref).max() > 0
    e_ = rel_l2(r["traces"][1:case.nft], ref)
    log(f"order-4 PZT plate {kernel} {precision}: {e_:.2e}")
    assert e_ < tol, e_


@pytest.mark.parametrize("kernel", [(1, 4), (1, 8), (3, 4), (3, 8)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_solid_order5_quasi_isotropic_plate_matches_oracle(kernel, precision, tol):
    """The 250 x 250 x 2 workload's element (order 5, 16-ply [0/45/-45/90]2s, 4 mm plate of 4 mm
    elements), actuator (75 kHz) and surface w sensors with two snapshots, on 16 x 16 x 2."""
    case = E.solid_plate_250x250_pzt(lx=0.064, ly=0.064, nex=16, ney=16, nft=200, force_xy=(0.032, 0.032),
                                     sensors=((0.044, 0.032), (0.032, 0.048)), n_frames=2)
    s, prep = E.solid_solver(case, kernel=kernel)
    r = s.run(case.dt, 2, case.nft, precision)
    s.close()
    states, _ = cached("solid_o5", lambda: _solid_oracle(case, prep, case.nft))
    outs = np.asarray(prep["outputs"])
    ref = states[2:case.nft + 1][:, outs]
    assert np.abs(ref).max() > 0
    e_tr = rel_l2(r["traces"][1:case.nft], ref)
    e_fr = max(rel_l2(r["frames"][i], states[nn]) for i, nn in enumerate(prep["frame_no"]))
    log(f"order-5 plate {kernel} {precision}: traces {e_tr:.2e}, snapshots {e_fr:.2e}")
    assert e_tr < tol and e_fr < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_solid_order4_pzt_voltage_long_run_matches_oracle(precision, tol):
    """The 100 x 100 x 2 workload's element, laminate, actuator and PZT voltage sensors over
    4000 steps on a 6 x 6 x 2 plate (30 mm): the whole 100 kHz burst (2500 steps) and its
    reflections, the voltages against the oracle's in both forms.  The autotuner's pick for
    the workload, the line kernel with 8-element patches."""
    case = E.solid_plate_100x100_pzt(lx=0.03, ly=0.03, nex=6, ney=6, nft=4000, force_xy=(0.015, 0.015),
                                     sensors=((0.025, 0.015), (0.015, 0.025)))
    s, prep = E.solid_solver(case, kernel=(1, 8))
    r = s.run(case.dt, 2, case.nft, precision, want_frames=False)
    s.close()

    def oracle():
        model, f = _solid_oracle_model(case, prep)
        t0 = time.perf_counter()
        o = model.run(case.dt, prep["st"], f, 2, case.nft, out_weights=_voltage_weights(case, prep))
        log(f"oracle 6x6x2 order 4, {case.nft} steps: {time.perf_counter() - t0:.1f} s")
        return o["traces"]

    ref = cached("solid_o4_long", oracle)[1:case.nft]
    assert np.abs(ref).max() > 0
    e_ = rel_l2(r["traces"][1:case.nft], ref)
    log(f"order-4 PZT plate, {case.nft} steps, {precision}: voltages {e_:.2e}")
    assert e_ < tol, e_


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32_split", 1e-3)])
def test_solid_order5_plate_long_run_matches_oracle(precision, tol):
    """The 250 x 250 x 2 workload's element, laminate, actuator and time step over 4000 steps
    on a 6 x 6 x 2 plate: the whole 75 kHz burst (3333 steps) and the reflections of a 24 mm
    plate, so that the FP32 form's error is stated against the FP64 oracle over a run long
    enough for it to accumulate.  Staged kernel, 8-element patches (the workload's pick)."""
    case = E.solid_plate_250x250_pzt(lx=0.024, ly=0.024, nex=6, ney=6, nft=4000, force_xy=(0.012, 0.012),
                                     sensors=((0.020, 0.012), (0.012, 0.004)), n_frames=0)
    s, prep = E.solid_solver(case, kernel=(3, 8))
    r = s.run(case.dt, 2, case.nft, precision, want_frames=False)
    s.close()

    def oracle():
        model, f = _solid_oracle_model(case, prep)
        t0 = time.perf_counter()
        o = model.run(case.dt, prep["st"], f, 2, case.nft, outputs=prep["outputs"])
        log(f"oracle 6x6x2 order 5, {case.nft} steps: {time.perf_counter() - t0:.1f} s")
        return o["traces"]

    ref = cached("solid_o5_long", oracle)[1:case.nft]
    assert np.abs(ref).max() > 0
    e_ = rel_l2(r["traces"][1:case.nft], ref)
    e_half = rel_l2(r["traces"][1:case.nft // 2], ref[:case.nft // 2 - 1])
    log(f"order-5 plate, {case.nft} steps, {precision}: traces {e_:.2e} (first half {e_half:.2e})")
    assert e_ < tol, e_


def _full_size_properties(case, kernel, steps):
    t0 = time.perf_counter()
    s, prep = E.solid_solver(case, kernel=kernel)
    log(f"set-up {case.nex}x{case.ney}x{case.nlay_el} order {case.order}: {time.perf_counter() - t0:.1f} s")
    last = 1 + steps
    try:
        r1 = s.run(case.dt, 2, last, "fp64", want_frames=False)["traces"]
        s.set_state(None, None)
        r2 = s.run(case.dt, 2, last, "fp64", want_frames=False)["traces"]
        s.set_state(None, None)
        r32 = s.run(case.dt, 2, last, "fp32_split", want_frames=False)["traces"]
        s.set_excitation(prep["ex_dofs"], 2.0 * np.asarray(prep["ex_amp"]), prep["st"])
        s.set_state(None, None)
        r_2x = s.run(case.dt, 2, last, "fp64", want_frames=False)["traces"]
    finally:
        s.close()
    assert np.isfinite(r1).all() and np.abs(r1).max() > 0
    assert np.array_equal(r1, r2)                      # rerun bitwise
    e_lin = rel_l2(r_2x, 2.0 * r1)                     # linear in the drive
    e_32 = rel_l2(r32, r1)
    log(f"full size: linearity {e_lin:.2e}, FP32 form {e_32:.2e}")
    assert e_lin < 1e-12 and e_32 < 1e-3


def test_solid_100x100x2_order4_full_size_properties():
    case = E.solid_plate_100x100_pzt(nft=400)
    case.sensors = tuple(case.sensors) + ((0.27, 0.25),)
    _full_size_properties(case, (1, 4), 300)


def test_solid_250x250x2_order5_full_size_properties():
    case = E.solid_plate_250x250_pzt(nft=400, n_frames=0)
    case.sensors = tuple(case.sensors) + ((0.51, 0.5),)
    _full_size_properties(case, (1, 8), 300)


# ---------------------------------------------------------------------------- damage sweep
def test_damage_sweep_cases_60_to_120_match_single_runs_and_oracle():
    cfg = B.damage_sweep(n_cases=3, seed=3, nft=400, sensors=((0.26, 0.25), (0.25, 0.27)))
    for c, n in zip(cfg.cases, (60, 90, 120)):
        c.n = n
    prep = B.prepare_batch(cfg)
    steps = 300
    s, _ = B.batch_solver(cfg, prep)
    r64 = s.run(cfg.dt, 1, steps, "fp64")["traces"][:steps]
    s.set_state(None, None)
    r32 = s.run(cfg.dt, 1, steps, "fp32_split")["traces"][:steps]
    s.close()
    per64, per32 = B.split(r64, cfg), B.split(r32, cfg)
    for k in range(3):
        cs = B.case_solver(cfg, prep, k)
        single = cs.run(cfg.dt, 1, steps, "fp64")["traces"][:steps]
        cs.close()
        assert np.abs(single).max() > 0
        e_single, e_32 = rel_l2(per64[k], single), rel_l2(per32[k], per64[k])
        log(f"case {k} ({cfg.cases[k].n}^2): batch vs single {e_single:.2e}, FP32 {e_32:.2e}")
        assert e_single < 1e-12 and e_32 < 1e-3
    nodes, coords = prep["meshes"][0]
    model = osh.ShellModel(nodes, coords, cfg.order, np.atleast_2d(prep["coef"]), stiff_scale=prep["scales"][0])
    f = np.zeros(model.dof)
    f[5 * prep["src"][0] + 4] = 1.0
    t0 = time.perf_counter()
    o = model.run(cfg.dt, prep["st"], f, 1, steps, [5 * j + 4 for j in prep["sens"][0]], cfg.out_quantity)
    log(f"oracle 60x60, {steps} steps: {time.perf_counter() - t0:.1f} s")
    e_o = rel_l2(per64[0], o["traces"][:steps])
    log(f"case 0 vs oracle: {e_o:.2e}")
    assert e_o < 1e-10


# ---------------------------------------------------------------------------- adjoint
def test_inversion_120x120_adjoint_matches_central_differences():
    case = I.inversion_120x120()  # the workload's own 15000 steps: the wave crosses the blobs
    prep = I.prepare(case)
    n_el = prep["nodes"].shape[0]
    s = I.make_solver(case, prep)
    try:
        obs, _ = I.traces(s, case, prep, prep["s_true"])
        one = np.ones(n_el)
        tr, _ = I.traces(s, case, prep, one, one)
        dJ = tr - obs
        # a misfit well above rounding: central differences of a misfit near zero measure
        # only their own O(h^2) term
        assert (dJ ** 2).sum() > 1e-4 * (obs ** 2).sum()
        s.set_material(prep["coef"], stiff_scale=one, rho_scale=one)
        s.set_state(None, None)
        g = s.adjoint(case.dt, 1, case.nft, dJ)

        def J(st, rh):
            t, _ = I.traces(s, case, prep, st, rh)
            return 0.5 * float(((t - obs) ** 2).sum())

        rng = np.random.default_rng(11)
        h = 1e-3
        for name, grad in (("stiffness", g["g_stiff"]), ("density", g["g_rho"])):
            top = np.argsort(-np.abs(grad))[:3]
            els = np.concatenate([top, rng.choice(n_el, 2, replace=False)])
            fd = []
            for el in els:
                v = []
                for sg in (1, -1):
                    st, rh = one.copy(), one.copy()
                    (st if name == "stiffness" else rh)[el] += sg * h
                    v.append(J(st, rh))
                fd.append((v[0] - v[1]) / (2 * h))
            fd = np.array(fd)
            err = np.abs(fd - grad[els]).max() / np.abs(grad[els]).max()
            log(f"120x120 adjoint {name}: max |fd - adjoint| / max |adjoint| = {err:.2e} on {els.tolist()}")
            assert np.abs(grad[els]).max() > 0 and err < 1e-4
    finally:
        s.close()
