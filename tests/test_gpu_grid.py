"""Wavefield on a uniform grid on the GPU (sem_grid_eval through the C-ABI) against the
oracle's spec2meshgrid restatement: every quantity 1-9, both surfaces, delaminated meshes
with the twin overwrite, and the frame files an engine run writes."""
import numpy as np
import pytest

import semhip
from semhip import engine as E, grid as G, matio
from cases import rel_l2
from oracle import mesh as om, grid as og

pytestmark = pytest.mark.gpu


def _mesh(order=4, perturb=0.2, seed=4):
    q, p = om.rect_quad_mesh(0.06, 0.05, 6, 5)
    rng = np.random.default_rng(seed)
    inner = (p[:, 0] > 0) & (p[:, 0] < 0.06) & (p[:, 1] > 0) & (p[:, 1] < 0.05)
    p = p.copy()
    p[inner, :2] += perturb * rng.uniform(-1, 1, (inner.sum(), 2)) * 0.01
    return om.quad2spectral(q, p, order)


@pytest.mark.parametrize("motion", range(1, 10))
@pytest.mark.parametrize("surface", ["top", "bottom"])
def test_eval_matches_oracle(motion, surface):
    nodes, coords = _mesh()
    frames = np.random.default_rng(motion).standard_normal((3, 5 * coords.shape[0]))
    o = og.spec2meshgrid(nodes, coords, 5, frames, 37, 29, motion, surface, h=2e-3)
    with G.WavefieldGrid(nodes, coords, 5, 37, 29) as g:
        d = g.eval(frames, motion, surface, h=2e-3)
        pts = g.points()
    assert np.array_equal(pts["owner"], o["owner"])
    assert d.shape == (3, 29, 37)
    assert rel_l2(d, o["data"]) < 1e-12


def test_delaminated_mesh_top_surface_uses_the_twins():
    nodes, coords = _mesh(perturb=0.0)
    n2, c2, du, da = semhip.split_delam(nodes, coords, [8, 9, 14, 15], 5)
    frames = np.random.default_rng(0).standard_normal((2, 5 * c2.shape[0]))
    kw = dict(n_search=nodes.shape[0], den_under=du, den_above=da)
    for surface in ("top", "bottom"):
        o = og.spec2meshgrid(n2, c2, 5, frames, 33, 33, 6, surface, h=2e-3, **kw)
        mesh = dict(nodes=n2, coords=c2, den_under=[du], den_above=[da])
        D = G.spec2meshgrid_flat_shell(1, mesh, 33, 33, "velocity", 6, surface, 2e-3, frames=frames)
        assert D.shape == (33, 33, 2)
        assert rel_l2(np.transpose(D, (2, 0, 1)), o["data"]) < 1e-12
    # the twins matter: without the overwrite the top surface differs inside the delamination
    with G.WavefieldGrid(n2, c2, 5, 33, 33, n_search=nodes.shape[0]) as g:
        plain = g.eval(frames, 6, "top", h=2e-3)
    o_top = og.spec2meshgrid(n2, c2, 5, frames, 33, 33, 6, "top", h=2e-3, **kw)
    assert rel_l2(plain, o_top["data"]) > 1e-3


def test_from_engine_frame_files(tmp_path):
    """Frames written by an engine run, read back from the files, interpolated."""
    case = E.shell_plate_40x40(lx=0.1, ly=0.1, nex=8, ney=8, order=5, force_xy=(0.05, 0.05), nft=300, n_frames=3,
                     sensors=((0.08, 0.05),), dt=1e-7)
    out = E.main_flat_shell_force_excit_c(case, output_name=tmp_path)
    prep = E.prepare_flat_shell(case)
    mesh = dict(nodes=prep["nodes"], coords=prep["coords"])
    D = G.spec2meshgrid_flat_shell(1, mesh, 41, 41, "velocity", 3, "top", prep["h"], input_name=tmp_path,
                                   n_frames=3, output_name=tmp_path)
    o = og.spec2meshgrid(prep["nodes"], prep["coords"], 6, out["frames"], 41, 41, 3, "top", h=prep["h"])
    assert np.abs(o["data"]).max() > 0
    assert rel_l2(np.transpose(D, (2, 0, 1)), o["data"]) < 1e-12
    import scipy.io as sio
    saved = sio.loadmat(str(tmp_path / "flat_shell_Vz_1_41x41top.mat"))["Data"]
    assert np.array_equal(saved, D)
