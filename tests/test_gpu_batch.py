"""Batched form: cases of uneven size stepped as one mesh (sem_upload_batch).

Bar: every case's FP64 batch traces equal its own single-case FP64 run to 1e-12 relative
L2 (only the basket summation order differs) and the oracle to 1e-10; the FP32-element
batch stays within 1e-3 of the FP64 batch for every case.
"""
import numpy as np
import pytest

from semhip import batch as B
from cases import rel_l2
from oracle import shell as osh

pytestmark = pytest.mark.gpu


def _cfg():
    cases = [B.DamageCase(n=4, centre=(0.02, 0.03), semi=(0.01, 0.006), angle=0.3, scale=0.6),
             B.DamageCase(n=6, centre=(0.04, 0.02), semi=(0.012, 0.008), angle=1.1, scale=0.8),
             B.DamageCase(n=5, centre=(0.03, 0.03), semi=(0.02, 0.02), angle=0.0, scale=0.5)]
    return B.BatchConfig(cases=cases, lx=0.06, ly=0.06, nft=300, dt=1e-7,
                         sensors=((0.045, 0.03), (0.03, 0.045)))


def test_batch_matches_single_case_runs_and_oracle():
    cfg = _cfg()
    prep = B.prepare_batch(cfg)
    s, off = B.batch_solver(cfg, prep)
    r64 = B.split(s.run(cfg.dt, 1, cfg.nft)["traces"][:cfg.nft], cfg)
    s.set_state(None, None)
    r32 = B.split(s.run(cfg.dt, 1, cfg.nft, "fp32_split")["traces"][:cfg.nft], cfg)
    s.close()
    s, off = B.batch_solver(cfg, prep, kernel=(2, 16))
    rmf = B.split(s.run(cfg.dt, 1, cfg.nft, "fp32_split")["traces"][:cfg.nft], cfg)
    s.close()
    s, off = B.batch_solver(cfg, prep, kernel=(1, 16))
    rln = B.split(s.run(cfg.dt, 1, cfg.nft)["traces"][:cfg.nft], cfg)
    s.close()
    s, off = B.batch_solver(cfg, prep, kernel=(0, 16))
    rpt = B.split(s.run(cfg.dt, 1, cfg.nft)["traces"][:cfg.nft], cfg)
    s.close()
    for k in range(len(cfg.cases)):
        assert rel_l2(rmf[k], r64[k]) < 1e-3, k
        assert rel_l2(rln[k], r64[k]) < 1e-12, k
        assert rel_l2(rpt[k], r64[k]) < 1e-12, k
    for k in range(len(cfg.cases)):
        assert np.abs(r64[k]).max() > 0
        with B.case_solver(cfg, prep, k) as c:
            single = c.run(cfg.dt, 1, cfg.nft)["traces"][:cfg.nft]
        assert rel_l2(r64[k], single) < 1e-12, k
        assert rel_l2(r32[k], r64[k]) < 1e-3, k
    nodes, coords = prep["meshes"][1]
    model = osh.ShellModel(nodes, coords, cfg.order, np.atleast_2d(prep["coef"]), stiff_scale=prep["scales"][1])
    f = np.zeros(model.dof)
    f[5 * prep["src"][1] + 4] = 1.0
    o = model.run(cfg.dt, prep["st"], f, 1, cfg.nft, [5 * j + 4 for j in prep["sens"][1]], cfg.out_quantity)
    assert rel_l2(r64[1], o["traces"][:cfg.nft]) < 1e-10
