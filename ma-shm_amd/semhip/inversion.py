"""Adjoint inversion workload: the misfit between sensor signals of a trial stiffness field
and those of a true (damaged) field, and its gradient with respect to the per-element
stiffness and density scales from the library's discrete adjoint (sem_adjoint).

True field: 1 minus 2-4 Gaussian blobs (amplitude U(0.1, 0.4), sigma U(10, 30) mm) drawn
from a recorded seed; trial field: 1 everywhere.  Loss J = 1/2 sum (tr - tr_obs)^2 over
9 point sensors (w-velocity) on a 3 x 3 grid and every step; dJ/dtr = tr - tr_obs.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import Solver, SEM_SHELL5, rect_quads, quad2spectral, laminate_shell, lamina_q
from .engine import hanning_signal, nearest_node


@dataclass
class InversionCase:
    n: int = 120
    lx: float = 0.5
    order: int = 5
    dt: float = 2e-8
    nft: int = 15000
    f_2: float = 50e3
    seed: int = 0


def inversion_120x120(**kw):
    """120 x 120 order-5 shell plate, 15000 steps of 2e-8 s, 9 sensors, Gaussian-blob
    stiffness field (InversionCase defaults)."""
    return InversionCase(**kw)


def true_field(centroids, seed):
    rng = np.random.default_rng(seed)
    s = np.ones(centroids.shape[0])
    for _ in range(int(rng.integers(2, 5))):
        c = rng.uniform(0.05, 0.45, 2)
        amp, sig = rng.uniform(0.1, 0.4), rng.uniform(10e-3, 30e-3)
        s -= amp * np.exp(-((centroids[:, 0] - c[0]) ** 2 + (centroids[:, 1] - c[1]) ** 2) / (2 * sig ** 2))
    return s


def prepare(case: InversionCase):
    quads, pos = rect_quads(case.lx, case.lx, case.n, case.n)
    nodes, coords = quad2spectral(quads, pos, case.order)
    q = lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    coef = laminate_shell(np.full(8, 0.25e-3), np.array([0, 90, 0, 90, 90, 0, 90, 0], float), 1570.0, q)
    cen = coords[nodes].mean(axis=1)
    g = [0.125, 0.25, 0.375]
    sens = [5 * nearest_node(coords, (x, y)) + 4 for y in g for x in g]
    src = 5 * nearest_node(coords, (case.lx / 2, case.lx / 2)) + 4
    _, st = hanning_signal(case.dt, case.nft, case.f_2 / 5, case.f_2, 0.0)
    return dict(nodes=nodes, coords=coords, coef=coef, s_true=true_field(cen, case.seed), sens=sens, src=src, st=st)


def make_solver(case, prep, device=0, kernel=(1, 8)):
    """Solver for the inversion plate; kernel = (variant, patch elements), by default the
    line-wise shell kernel with 8-element patches (the autotuner's pick for a single plate
    in both forms)."""
    s = Solver(device)
    if kernel is not None:
        s.set_kernel(*kernel)
    s.upload_mesh(SEM_SHELL5, prep["nodes"], prep["coords"], case.order + 1)
    s.set_excitation([prep["src"]], [1.0], prep["st"])
    s.set_outputs(prep["sens"], "vel")
    return s


def traces(s, case, prep, stiff, rho=None, precision="fp64"):
    s.set_material(prep["coef"], stiff_scale=stiff, rho_scale=rho)
    s.set_state(None, None)
    r = s.run(case.dt, 1, case.nft, precision, want_frames=False)
    if r["status"] != "ok":
        raise RuntimeError("forward run diverged")
    return r["traces"][:case.nft], r["seconds"]


def adjoint_sweeps(case: InversionCase, device=0, log=print, kernel=(1, 8), seed=1):
    """The checkpointed forward and reverse sweeps alone (sem_adjoint, which launches its
    kernels one by one), in FP64 and in the FP32 form, at the trial field with a seeded
    synthetic misfit derivative dJ/dtraces: no forward run of the signals, so the kernels of
    the adjoint are the only work.  Returns per form the device seconds of the sweeps and the
    relative L2 distance of the FP32-form gradients from the FP64 ones."""
    prep = prepare(case)
    n_el = prep["nodes"].shape[0]
    s = make_solver(case, prep, device, kernel)
    dJ = np.random.default_rng(seed).standard_normal((case.nft, len(prep["sens"])))
    out = {"dof": int(5 * prep["coords"].shape[0]), "elements": n_el, "steps": case.nft, "dJ_seed": seed}
    grads = {}
    for prec in ("fp64", "fp32_split"):
        s.set_material(prep["coef"], stiff_scale=np.ones(n_el), rho_scale=np.ones(n_el))
        s.set_state(None, None)
        g = s.adjoint(case.dt, 1, case.nft, dJ, precision=prec)
        grads[prec] = g
        out[prec] = {"adjoint_seconds": g["seconds"]}
        log(f"[adjoint sweeps] {prec}: forward+reverse sweeps {g['seconds']:.3f} s over {case.nft} steps")
    s.close()
    for k in ("g_stiff", "g_rho"):
        ref = grads["fp64"][k]
        out["fp32_split"][f"rel_l2_vs_fp64_{k}"] = float(np.linalg.norm(grads["fp32_split"][k] - ref) /
                                                         np.linalg.norm(ref))
    return out


def gradient_check(case: InversionCase, n_fd=20, h=1e-3, device=0, log=print, kernel=(1, 8)):
    """Observed signals from the true field (FP64), gradients of J at the trial field by
    the adjoint in FP64 and in the FP32 form, and central differences of J (FP64 runs) on
    n_fd random elements (none when n_fd is 0).  Returns a dict of timings and errors."""
    prep = prepare(case)
    n_el = prep["nodes"].shape[0]
    s = make_solver(case, prep, device, kernel)
    obs, _ = traces(s, case, prep, prep["s_true"])
    s0, r0 = np.ones(n_el), np.ones(n_el)
    out = {"dof": int(5 * prep["coords"].shape[0]), "elements": n_el, "steps": case.nft, "seed": case.seed}
    grads = {}
    for prec in ("fp64", "fp32_split"):
        tr, fwd_s = traces(s, case, prep, s0, r0, prec)
        dJ = tr - obs
        s.set_material(prep["coef"], stiff_scale=s0, rho_scale=r0)
        s.set_state(None, None)
        t0 = time.perf_counter()
        g = s.adjoint(case.dt, 1, case.nft, dJ, precision=prec)
        wall = time.perf_counter() - t0
        grads[prec] = g
        out[prec] = {"forward_seconds": fwd_s, "adjoint_seconds": g["seconds"], "adjoint_wall": wall,
                     "J": 0.5 * float((dJ ** 2).sum())}
        log(f"[adjoint] {prec}: forward {fwd_s:.3f} s, forward+reverse sweeps {g['seconds']:.3f} s")
    gs64 = grads["fp64"]["g_stiff"]
    for prec in ("fp32_split",):
        out[prec]["rel_l2_vs_fp64"] = {
            "g_stiff": float(np.linalg.norm(grads[prec]["g_stiff"] - gs64) / np.linalg.norm(gs64)),
            "g_rho": float(np.linalg.norm(grads[prec]["g_rho"] - grads["fp64"]["g_rho"]) /
                           np.linalg.norm(grads["fp64"]["g_rho"]))}
    if n_fd <= 0:
        s.close()
        out["fd"] = None
        return out
    rng = np.random.default_rng(case.seed + 1)
    els = rng.choice(n_el, n_fd, replace=False)
    fd_s, fd_r = [], []
    t0 = time.perf_counter()

    def J(st, rh):
        tr, _ = traces(s, case, prep, st, rh)
        return 0.5 * float(((tr - obs) ** 2).sum())

    for e in els:
        for which, acc in (("s", fd_s), ("r", fd_r)):
            v = []
            for sg in (1, -1):
                st, rh = s0.copy(), r0.copy()
                (st if which == "s" else rh)[e] += sg * h
                v.append(J(st, rh))
            acc.append((v[0] - v[1]) / (2 * h))
    s.close()
    fd_s, fd_r = np.array(fd_s), np.array(fd_r)
    ad_s, ad_r = gs64[els], grads["fp64"]["g_rho"][els]
    out["fd"] = {"elements": [int(x) for x in els], "h": h, "seconds": time.perf_counter() - t0,
                 "max_rel_err_stiff": float(np.abs(fd_s - ad_s).max() / np.abs(ad_s).max()),
                 "max_rel_err_rho": float(np.abs(fd_r - ad_r).max() / np.abs(ad_r).max()),
                 "fp32_max_rel_err_stiff": float(np.abs(fd_s - grads["fp32_split"]["g_stiff"][els]).max() /
                                                 np.abs(ad_s).max())}
    return out
