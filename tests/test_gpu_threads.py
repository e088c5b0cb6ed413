"""Separate contexts on separate host threads at once: a flat-shell run and a 3D-solid run
(FP64, 8-element order-5 patches, whose kernel needs the >64 KB LDS opt-in) started together
on two threads give bitwise the traces each gives when run alone.  The binding's calls
release the interpreter lock, so the two runs overlap on the host and on the device."""
import threading

import numpy as np
import pytest

import semhip
from cases import shell_pair, solid_pair
from oracle.common import hanning_signal

pytestmark = pytest.mark.gpu


def _shell_case():
    model, inp = shell_pair(nex=6, ney=5, order=5)
    dt, nft = 1e-7, 300
    _, st = hanning_signal(dt, nft, 50e3 / 5, 50e3, 0.0)
    dof = model.dof
    src = dof // 2 - (dof // 2) % 5 + 2
    outs = [2, src, dof - 3]

    def go():
        with semhip.Solver(0) as s:
            s.set_kernel(1, 8)
            s.upload_mesh(semhip.SEM_SHELL5, inp["nodes"], inp["coords"], inp["n_xy"])
            s.set_material(inp["coef"])
            s.set_excitation([src], [1.0], st)
            s.set_outputs(outs, "vel")
            return s.run(dt, 1, nft)["traces"]
    return go


def _solid_case():
    model, inp = solid_pair(nex=4, ney=4, nlay=2, order=5, lx=0.02, ly=0.02)
    dt, nft = 2e-8, 200
    _, st = hanning_signal(dt, nft, 100e3 / 5, 100e3, 0.0)
    src = 3 * (model.n_nodes // 2) + 2
    outs = [3 * (model.n_nodes // 3) + 2, src]

    def go():
        with semhip.Solver(0) as s:
            s.set_kernel(1, 8)
            s.upload_mesh(semhip.SEM_SOLID3, inp["nodes"], inp["coords"], inp["n_xy"], inp["n_z"])
            s.set_material(inp["coef"])
            s.set_excitation([src], [1.0], st)
            s.set_outputs(outs, "disp_post")
            return s.run(dt, 2, nft)["traces"]
    return go


def test_two_contexts_on_two_threads():
    cases = [_shell_case(), _solid_case()]
    together = [None, None]
    errors = []

    def worker(k):
        try:
            together[k] = cases[k]()
        except Exception as exc:  # reported below: an exception in a thread does not fail the test
            errors.append(exc)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a run did not finish"
    assert not errors, errors
    alone = [c() for c in cases]
    for a, b in zip(together, alone):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b)
