"""Error paths of the C-ABI (through the Python binding): calls out of order, indices out
of range, unsupported element kernels and orders, inverted elements, signals shorter than
the run.  Every call is refused on the host with its code and message, nothing reaches a
kernel, and the context stays usable: a valid call afterwards gives the same result as a
fresh context."""
import numpy as np
import pytest

import semhip
from cases import rel_l2, shell_pair, solid_pair
from oracle.common import hanning_signal

pytestmark = pytest.mark.gpu

E_ARG, E_STATE, E_UNSUPPORTED = -1, -3, -6


def _shell(kernel=None):
    model, inp = shell_pair(nex=3, ney=2, order=4)
    s = semhip.Solver(0)
    if kernel is not None:
        s.set_kernel(*kernel)
    s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], inp["n_xy"])
    s.set_material(inp["coef"])
    return s, model, inp


def _code(exc):
    return exc.value.code


def test_calls_out_of_order_are_state_errors():
    with semhip.Solver(0) as s:
        with pytest.raises(semhip.SemError) as e:
            s.run(1e-7, 1, 10)
        assert _code(e) == E_STATE
        with pytest.raises(semhip.SemError) as e:
            s.set_excitation([0], [1.0], np.ones(10))
        assert _code(e) == E_STATE and "mesh" in str(e.value)


def test_mesh_with_a_node_id_out_of_range_is_refused():
    _, inp = shell_pair(nex=2, ney=2, order=4)
    nodes = inp["nodes"].copy()
    nodes[1, 7] = inp["coords"].shape[0]  # one past the last node
    with semhip.Solver(0) as s:
        with pytest.raises(semhip.SemError) as e:
            s.upload_mesh(semhip.SEM_SHELL5, nodes, inp["coords"], inp["n_xy"])
        assert _code(e) == E_ARG and "out of range" in str(e.value)


def test_unsupported_orders_are_refused():
    _, inp = shell_pair(nex=2, ney=2, order=4)
    with semhip.Solver(0) as s:
        with pytest.raises(semhip.SemError) as e:
            s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], 9)  # n_xy = 9
        assert _code(e) in (E_ARG, E_UNSUPPORTED)


def test_inverted_element_is_refused():
    """Mirroring one element's local numbering along xi makes its Jacobian negative."""
    _, inp = shell_pair(nex=2, ney=2, order=4)
    n = inp["n_xy"]
    nodes = inp["nodes"].copy()
    nodes[0] = nodes[0].reshape(n, n)[:, ::-1].reshape(-1)
    with semhip.Solver(0) as s:
        with pytest.raises(semhip.SemError) as e:
            s.upload_mesh(semhip.SEM_SHELL5, nodes, inp["coords"], n)
        assert _code(e) == E_ARG and "Jacobian" in str(e.value)


def test_bad_kernel_choices_are_refused():
    with semhip.Solver(0) as s:
        with pytest.raises(semhip.SemError) as e:
            s.set_kernel(4, 0)
        assert _code(e) == E_ARG
    model, inp_shell = shell_pair(nex=2, ney=2, order=4)
    with semhip.Solver(0) as s:
        s.set_kernel(3, 8)  # the staged kernel is for solids
        with pytest.raises(semhip.SemError) as e:
            s.upload_mesh(semhip.SEM_SHELL5, inp_shell["nodes"], inp_shell["coords"], inp_shell["n_xy"])
        assert _code(e) == E_UNSUPPORTED
    _, inp = solid_pair(nex=2, ney=2, nlay=2, order=4)
    with semhip.Solver(0) as s:
        s.set_kernel(1, 16)  # the line kernel takes 4- or 8-element solid patches
        with pytest.raises(semhip.SemError) as e:
            s.upload_mesh(semhip.SEM_SOLID3, inp["nodes"], inp["coords"], inp["n_xy"], inp["n_z"])
        assert _code(e) == E_UNSUPPORTED and "patches" in str(e.value)


def test_bad_excitation_outputs_and_frames_leave_the_context_usable():
    dt, nft = 1e-7, 60
    _, st = hanning_signal(dt, nft, 50e3 / 5, 50e3, 0.0)
    s, model, _ = _shell(kernel=(1, 8))
    dof = model.dof
    with pytest.raises(semhip.SemError) as e:
        s.set_excitation([dof], [1.0], st)  # one past the last DOF
    assert _code(e) == E_ARG and "excitation" in str(e.value)
    with pytest.raises(semhip.SemError) as e:
        s.set_outputs([-1], "vel")
    assert _code(e) == E_ARG
    with pytest.raises(semhip.SemError) as e:
        s.set_frames([20, 10])
    assert _code(e) == E_ARG and "ascend" in str(e.value)
    # the same context after the refusals against a fresh one: bitwise equal
    src, sens = dof // 2 - (dof // 2) % 5 + 4, [4, dof - 1]
    s.set_excitation([src], [1.0], st)
    s.set_outputs(sens, "vel")
    r1 = s.run(dt, 1, nft)["traces"]
    s.close()
    s2, _, _ = _shell(kernel=(1, 8))
    s2.set_excitation([src], [1.0], st)
    s2.set_outputs(sens, "vel")
    r2 = s2.run(dt, 1, nft)["traces"]
    s2.close()
    assert np.abs(r1).max() > 0
    assert np.array_equal(r1, r2)


def test_signal_shorter_than_the_run_is_refused():
    dt = 1e-7
    _, st = hanning_signal(dt, 40, 50e3 / 5, 50e3, 0.0)
    s, model, _ = _shell()
    s.set_excitation([4], [1.0], st)
    s.set_outputs([4], "disp")
    with pytest.raises(semhip.SemError) as e:
        s.run(dt, 1, 41)
    assert _code(e) == E_ARG and "signal" in str(e.value)
    r = s.run(dt, 1, 40)  # the full signal length runs
    assert r["status"] == "ok" and r["nn_done"] == 40
    s.close()


def test_unknown_precision_is_refused():
    s, model, _ = _shell()
    with pytest.raises((semhip.SemError, KeyError, ValueError)):
        s.run(1e-7, 1, 10, precision="fp16")
    s.close()


def test_refused_calls_keep_the_previous_setting():
    """A refused set_outputs / set_frames / set_excitation leaves the previous outputs,
    frames and sources in place (library and binding agree on the counts), so a run
    afterwards equals a run that never saw the refused calls."""
    dt, nft = 1e-7, 40
    _, st = hanning_signal(dt, nft, 50e3 / 5, 50e3, 0.0)

    def configure():
        s, model, _ = _shell(kernel=(1, 8))
        dof = model.dof
        s.set_excitation([dof // 2 - (dof // 2) % 5 + 4], [1.0], st)
        s.set_outputs([4, 9, dof - 1], "vel")
        s.set_frames([10, 20])
        return s, dof

    s, dof = configure()
    with pytest.raises(semhip.SemError):
        s.set_outputs([dof], "vel")
    with pytest.raises(semhip.SemError):
        s.set_frames([30, 25])
    with pytest.raises(semhip.SemError):
        s.set_frames([0, 5])
    with pytest.raises(semhip.SemError):
        s.set_excitation([-5], [1.0], st)
    r1 = s.run(dt, 1, nft)
    s.close()
    s2, _ = configure()
    r2 = s2.run(dt, 1, nft)
    s2.close()
    assert r1["traces"].shape == (nft, 3) and r1["frames"].shape == (2, dof)
    assert np.abs(r1["traces"]).max() > 0
    assert np.array_equal(r1["traces"], r2["traces"])
    assert np.array_equal(r1["frames"], r2["frames"])


def test_refused_upload_leaves_no_mesh():
    """After a refused upload the context holds no mesh (and no material): calls that need
    one are state errors, and a valid upload afterwards works as on a fresh context."""
    s, model, inp = _shell(kernel=(1, 8))
    n = inp["n_xy"]
    bad = inp["nodes"].copy()
    bad[0] = bad[0].reshape(n, n)[:, ::-1].reshape(-1)
    with pytest.raises(semhip.SemError):
        s.upload_mesh(semhip.SEM_SHELL5, bad, inp["coords"], n)
    with pytest.raises(semhip.SemError) as e:
        s.set_material(inp["coef"])
    assert _code(e) == E_STATE
    with pytest.raises(semhip.SemError) as e:
        s.apply_K(np.zeros(model.dof))
    assert _code(e) == E_STATE
    s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], n)
    s.set_material(inp["coef"])
    u = np.random.default_rng(4).standard_normal(model.dof)
    f1 = s.apply_K(u)
    s.close()
    s2, _, _ = _shell(kernel=(1, 8))
    f2 = s2.apply_K(u)
    s2.close()
    assert np.array_equal(f1, f2)


def test_kernel_change_after_upload():
    """The variant takes effect at the next run, the patch size at the next upload: a variant
    that does not take the uploaded patches is a state error until the mesh is uploaded again."""
    s, model, inp = _shell(kernel=(1, 8))
    u = np.random.default_rng(5).standard_normal(model.dof)
    f8 = s.apply_K(u)
    s.set_kernel(1, 16)  # patch size pending: the mesh keeps its 8-element patches
    assert s.kernel() == (1, 8)
    assert np.array_equal(s.apply_K(u), f8)
    s.set_kernel(0, 16)  # the point-wise shell kernel takes 16-element patches only
    with pytest.raises(semhip.SemError) as e:
        s.apply_K(u)
    assert _code(e) == E_STATE and "upload the mesh again" in str(e.value)
    s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], inp["n_xy"])
    s.set_material(inp["coef"])
    assert s.kernel() == (0, 16)
    assert rel_l2(s.apply_K(u), f8) < 1e-13  # element kernels agree to rounding
    s.close()
