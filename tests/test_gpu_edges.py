"""Edge cases of the step path: the smallest meshes (one element: a single patch in which
one wavefront or thread group of the workgroup has an element and the others idle), every
element kernel on them, a one-step run, a run with no outputs, and the empty inputs the
host refuses (no elements, a step range that ends before it starts)."""
import numpy as np
import pytest

import semhip
from cases import rel_l2, shell_pair, solid_pair
from oracle.common import hanning_signal

pytestmark = pytest.mark.gpu

E_ARG, E_UNSUPPORTED = -1, -6

SHELL_KERNELS = [(0, 16), (1, 16), (1, 8), (2, 16)]
SOLID_KERNELS = [(0, 4), (1, 4), (1, 8), (3, 4), (3, 8)]


def _shell(inp, kernel):
    s = semhip.Solver(0)
    s.set_kernel(*kernel)
    s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], inp["n_xy"])
    s.set_material(inp["coef"])
    return s


def _solid(inp, kernel):
    s = semhip.Solver(0)
    s.set_kernel(*kernel)
    s.upload_mesh(semhip.SEM_SOLID3, inp["nodes"], inp["coords"], inp["n_xy"], inp["n_z"])
    s.set_material(inp["coef"])
    return s


@pytest.mark.parametrize("kernel", SHELL_KERNELS)
@pytest.mark.parametrize("order", [3, 5])
def test_single_shell_element(order, kernel):
    model, inp = shell_pair(nex=1, ney=1, order=order)
    u = np.random.default_rng(order).standard_normal(model.dof)
    ref = model.internal_force(u)
    with _shell(inp, kernel) as s:
        if kernel[0] == 2:
            # the matrix-core kernel runs the FP32 form of order-4 and order-5 shells only
            with pytest.raises(semhip.SemError) as e:
                s.apply_K(u)
            assert e.value.code == E_UNSUPPORTED
            if order != 5:
                with pytest.raises(semhip.SemError) as e:
                    s.apply_K(u, "fp32_split")
                assert e.value.code == E_UNSUPPORTED
                return
        else:
            assert rel_l2(s.apply_K(u), ref) < 1e-12
        assert rel_l2(s.apply_K(u, "fp32_split"), ref) < 1e-4


@pytest.mark.parametrize("kernel", SOLID_KERNELS)
@pytest.mark.parametrize("order,order_z", [(3, 3), (4, 4), (4, 2)])
def test_single_solid_element(order, order_z, kernel):
    model, inp = solid_pair(nex=1, ney=1, nlay=1, order=order, order_z=order_z)
    u = np.random.default_rng(order).standard_normal(model.dof)
    ref = model.internal_force(u)
    with _solid(inp, kernel) as s:
        assert rel_l2(s.apply_K(u), ref) < 1e-12
        assert rel_l2(s.apply_K(u, "fp32_split"), ref) < 1e-4


@pytest.mark.parametrize("kernel", [(0, 16), (1, 8)])
def test_single_shell_element_run(kernel):
    """A whole run on one element (every node on the plate edge, none shared with another
    patch) against the oracle's run, FP64."""
    model, inp = shell_pair(nex=1, ney=1, order=5)
    dt, nft = 1e-7, 200
    _, st = hanning_signal(dt, nft, 50e3 / 5, 50e3, 0.0)
    src, outs = 5 * 12 + 2, [2, 5 * 20 + 2, 5 * 35 + 2]
    f_ext = np.zeros(model.dof)
    f_ext[src] = 1.0
    o = model.run(dt, st, f_ext, 1, nft, outs, "vel", [], "vel")
    with _shell(inp, kernel) as s:
        s.set_excitation([src], [1.0], st)
        s.set_outputs(outs, "vel")
        r = s.run(dt, 1, nft)
    assert np.abs(o["traces"]).max() > 0
    assert rel_l2(r["traces"], o["traces"]) < 1e-10


def test_one_step_and_no_outputs():
    """nn_first == nn_last runs exactly one step and the next run continues from its state;
    with no outputs set a run returns an empty trace block.  Steps 1 and 2..nft without
    outputs leave bitwise the state of steps 1..nft with an output."""
    model, inp = shell_pair(nex=3, ney=2, order=4)
    dt, nft = 1e-7, 40
    _, st = hanning_signal(dt, nft, 50e3 / 5, 50e3, 0.0)
    src = 5 * (model.n_nodes // 2) + 2
    with _shell(inp, (1, 8)) as s:
        s.set_excitation([src], [1.0], st)
        r = s.run(dt, 1, 1)
        assert r["status"] == "ok" and r["nn_done"] == 1
        assert r["traces"].size == 0
        r = s.run(dt, 2, nft)
        assert r["nn_done"] == nft
        u1, uo1 = s.get_state()
    with _shell(inp, (1, 8)) as s:
        s.set_excitation([src], [1.0], st)
        s.set_outputs([src], "disp")
        r2 = s.run(dt, 1, nft)
        u2, uo2 = s.get_state()
    assert np.abs(r2["traces"]).max() > 0 and np.abs(u1).max() > 0
    assert np.array_equal(u1, u2) and np.array_equal(uo1, uo2)


def test_empty_inputs_are_refused():
    _, inp = shell_pair(nex=2, ney=2, order=4)
    with semhip.Solver(0) as s:
        with pytest.raises(semhip.SemError) as e:
            s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"][:0], inp["coords"], inp["n_xy"])
        assert e.value.code == E_ARG and "empty" in str(e.value)
    dt = 1e-7
    _, st = hanning_signal(dt, 20, 50e3 / 5, 50e3, 0.0)
    with _shell(inp, (1, 8)) as s:
        s.set_excitation([4], [1.0], st)
        with pytest.raises(semhip.SemError) as e:
            s.run(dt, 5, 4)
        assert e.value.code == E_ARG and "step range" in str(e.value)
        assert s.run(dt, 1, 20)["nn_done"] == 20


def test_runs_into_one_trace_buffer():
    """Consecutive runs writing into one caller-provided trace buffer (Solver.run traces_out)
    give, bit for bit, the traces of one run over the whole range; a buffer of the wrong shape
    or type is refused before the library is called."""
    model, inp = shell_pair(nex=3, ney=3, order=4)
    dt, nft = 1e-7, 90
    _, st = hanning_signal(dt, nft, 50e3 / 5, 50e3, 0.0)
    src = 5 * (model.dof // 10) + 4
    outs = [5 * 3 + 4, 5 * 7 + 4]
    with _shell(inp, (1, 8)) as s:
        s.set_excitation([src], [1.0], st)
        s.set_outputs(outs, "vel")
        whole = s.run(dt, 1, nft)["traces"]
        s.set_state(None, None)
        buf = np.full((2, nft), np.nan)
        for n0, n1 in ((1, 1), (2, 40), (41, 70), (71, nft)):
            r = s.run(dt, n0, n1, ld=nft, traces_out=buf)
            assert np.shares_memory(r["traces"], buf)
        assert np.array_equal(buf.T, whole)
        for bad in (np.zeros((2, nft - 1)), np.zeros((2, nft), np.float32), np.zeros((nft, 2)).T):
            with pytest.raises(ValueError):
                s.run(dt, 1, 5, ld=nft, traces_out=bad)
