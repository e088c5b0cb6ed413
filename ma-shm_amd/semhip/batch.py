"""Batched damage sweep: many flat-shell plates of uneven size, each with its own
delamination-proxy stiffness field, stepped together as one mesh (sem_upload_batch).

The cases share the time loop (one captured graph of steps for the whole batch) and the
element kernels; each keeps its own nodes, DOFs, source and sensors.  Results per case do
not depend on the other cases (the union has no coupling), so each case's FP64 batch
result is its own FP64 solve.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import Solver, SEM_SHELL5, rect_quads, quad2spectral, laminate_shell, lamina_q
from .engine import hanning_signal, nearest_node


@dataclass
class DamageCase:
    n: int                 # elements per side
    centre: tuple          # ellipse centre (m)
    semi: tuple            # semi-axes (m)
    angle: float           # rad
    scale: float           # stiffness scale inside the ellipse


@dataclass
class BatchConfig:
    cases: list
    lx: float = 0.5
    ly: float = 0.5
    order: int = 5
    lh: tuple = (0.25e-3,) * 8
    angles: tuple = (0, 90, 0, 90, 90, 0, 90, 0)
    E11: float = 130e9
    E22: float = 10e9
    G12: float = 5e9
    G13: float = 5e9
    G23: float = 3.5e9
    nu12: float = 0.3
    rho: float = 1570.0
    f_2: float = 50e3
    cycles: float = 5.0
    # one step for every case: 2e-8 s keeps the finest plate (120 order-5 elements per side,
    # 0.49 mm GLL spacing) stable; the 40 x 40 plate of the same laminate runs at 1e-7 s
    dt: float = 2e-8
    nft: int = 8000
    sensors: tuple = ((0.40, 0.25), (0.25, 0.40), (0.10, 0.25), (0.25, 0.10))
    out_quantity: str = "vel"
    seed: int = 0


def damage_sweep(n_cases=256, n_min=60, n_max=120, seed=0, **kw):
    """256 plates of 0.5 x 0.5 m, side n ~ U{n_min..n_max} elements of order 5; per case
    one ellipse (centre uniform in the plate, semi-axes U(5, 20) mm, angle U(0, pi)) whose
    elements have stiffness scaled by U(0.5, 0.9); centre 50 kHz burst; 8000 steps of
    2e-8 s."""
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n_cases):
        cases.append(DamageCase(n=int(rng.integers(n_min, n_max + 1)),
                                centre=(float(rng.uniform(0, 0.5)), float(rng.uniform(0, 0.5))),
                                semi=(float(rng.uniform(5e-3, 20e-3)), float(rng.uniform(5e-3, 20e-3))),
                                angle=float(rng.uniform(0, np.pi)), scale=float(rng.uniform(0.5, 0.9))))
    return BatchConfig(cases=cases, seed=seed, **kw)


def _stiffness_field(nodes, coords, c: DamageCase):
    cen = coords[nodes].mean(axis=1)
    dx, dy = cen[:, 0] - c.centre[0], cen[:, 1] - c.centre[1]
    ca, sa = np.cos(c.angle), np.sin(c.angle)
    u, v = ca * dx + sa * dy, -sa * dx + ca * dy
    inside = (u / c.semi[0]) ** 2 + (v / c.semi[1]) ** 2 <= 1.0
    s = np.ones(nodes.shape[0])
    s[inside] = c.scale
    return s


def prepare_batch(cfg: BatchConfig):
    q = lamina_q(cfg.E11, cfg.E22, cfg.G12, cfg.G13, cfg.G23, cfg.nu12)
    coef = laminate_shell(np.asarray(cfg.lh), np.asarray(cfg.angles, float), cfg.rho, q)
    meshes, scales, src, sens = [], [], [], []
    for c in cfg.cases:
        quads, pos = rect_quads(cfg.lx, cfg.ly, c.n, c.n)
        nodes, coords = quad2spectral(quads, pos, cfg.order)
        meshes.append((nodes, coords))
        scales.append(_stiffness_field(nodes, coords, c))
        src.append(nearest_node(coords, (cfg.lx / 2, cfg.ly / 2)))
        sens.append([nearest_node(coords, s) for s in cfg.sensors])
    _, st = hanning_signal(cfg.dt, cfg.nft, cfg.f_2 / cfg.cycles, cfg.f_2, 0.0)
    return dict(meshes=meshes, scales=scales, src=src, sens=sens, coef=coef, st=st,
                dof=int(sum(5 * m[1].shape[0] for m in meshes)),
                elements=int(sum(m[0].shape[0] for m in meshes)))


def batch_solver(cfg: BatchConfig, prep, device=0, kernel=(1, 8)):
    """One Solver holding every case.  Outputs: the cases' sensors in case order
    (len(cfg.sensors) per case).  kernel = (variant, patch elements), by default the
    line-wise kernel with 8-element patches; None selects the library default."""
    s = Solver(device)
    if kernel is not None:
        s.set_kernel(*kernel)
    off, _ = s.upload_batch(SEM_SHELL5, prep["meshes"], cfg.order + 1)
    s.set_material(prep["coef"], stiff_scale=np.concatenate(prep["scales"]))
    s.set_excitation([5 * (o + j) + 4 for o, j in zip(off, prep["src"])], 1.0, prep["st"])
    s.set_outputs([5 * (o + k) + 4 for o, ks in zip(off, prep["sens"]) for k in ks], cfg.out_quantity)
    return s, off


def case_solver(cfg: BatchConfig, prep, k, device=0, kernel=(1, 8)):
    """Solver of case k alone (the same inputs as its part of the batch)."""
    nodes, coords = prep["meshes"][k]
    s = Solver(device)
    if kernel is not None:
        s.set_kernel(*kernel)
    s.upload_mesh(SEM_SHELL5, nodes, coords, cfg.order + 1)
    s.set_material(prep["coef"], stiff_scale=prep["scales"][k])
    s.set_excitation([5 * prep["src"][k] + 4], 1.0, prep["st"])
    s.set_outputs([5 * j + 4 for j in prep["sens"][k]], cfg.out_quantity)
    return s


def split(traces, cfg: BatchConfig):
    """[steps x n_cases*n_sensors] -> list of [steps x n_sensors] per case."""
    m = len(cfg.sensors)
    return [traces[:, k * m:(k + 1) * m] for k in range(len(cfg.cases))]
