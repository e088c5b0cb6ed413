"""The reference's file formats (CPU): mesh files round-trip with 1-based ids on disk, and
engine outputs carry the reference's file and variable names and layouts."""
import numpy as np
import scipy.io as sio

import semhip
from semhip import matio


def test_mesh_file_round_trip(tmp_path):
    q, p = semhip.rect_quads(0.06, 0.05, 6, 5)
    nodes, coords = semhip.quad2spectral(q, p, 4)
    n2, c2, du, da = semhip.split_delam(nodes, coords, [8, 9, 14, 15], 5)
    IG, IL = semhip.baskets(n2, c2.shape[0])
    f = tmp_path / "plate.mat"
    matio.save_mesh(f, n2, c2, 4, I_G=IG, I_L=IL, pztEl=[np.array([0, 1])], delamEl=[du, da],
                    den_under=[du], den_above=[da])
    raw = sio.loadmat(str(f))
    assert raw["nodes"].min() == 1 and raw["nodes"].dtype == np.float64          # MATLAB layout
    assert raw["nodes"].shape == n2.shape and raw["coords"].shape == c2.shape
    m = matio.load_mesh(f)
    assert np.array_equal(m["nodes"], n2) and m["nodes"].dtype == np.int32
    assert np.array_equal(m["coords"], c2)
    assert m["shape_order"] == 4
    assert np.array_equal(m["I_G"], IG) and np.array_equal(m["I_L"], IL)
    assert np.array_equal(m["den_under"][0], du) and np.array_equal(m["den_above"][0], da)
    assert np.array_equal(m["pztEl"][0], [0, 1])
    assert np.array_equal(np.sort(np.concatenate(m["delamEl"])), np.sort(np.concatenate([du, da])))
    assert np.allclose(m["mesh_min"], c2.min(axis=0)) and np.allclose(m["mesh_max"], c2.max(axis=0))


def test_output_files_match_reference_names(tmp_path):
    nft, n_out, dof = 50, 4, 5 * 9
    displ = np.random.default_rng(0).standard_normal((nft, n_out))
    t = np.linspace(0, 1e-5, nft)
    matio.save_outputs(tmp_path, 3, displ=displ,
                       time=dict(minTime=1e-6, averageTime=2e-6, TotalNofNodes=9, t=t))
    d = sio.loadmat(str(tmp_path / "displ3.mat"))
    assert np.array_equal(d["displ"], displ)
    tm = sio.loadmat(str(tmp_path / "time3.mat"))
    assert {"minTime", "averageTime", "TotalNofNodes", "t"} <= set(tm)
    assert tm["TotalNofNodes"].item() == 9 and np.allclose(tm["t"].reshape(-1), t)
    frames = np.arange(2 * dof, dtype=float).reshape(2, dof)
    matio.save_frames(tmp_path, frames, "velocity")
    for c in (1, 2):
        for k, nm in enumerate(("x", "Fix", "y", "Fiy", "z")):
            v = sio.loadmat(str(tmp_path / f"V{nm}_frame{c}.mat"))[f"V{nm}"].reshape(-1)
            assert np.array_equal(v, frames[c - 1, k::5])


def test_engine_reads_a_delaminated_mesh_file(tmp_path):
    """The engine mirror set up from a mesh file equals the same plate built in memory."""
    from semhip import engine as E
    kw = dict(lx=0.1, ly=0.1, nex=8, ney=8, order=5, force_xy=(0.05, 0.05), nft=50, n_frames=0,
              sensors=((0.08, 0.05),), delamination_layer=3)
    a = E.prepare_flat_shell(E.shell_plate_40x40(delam_rect=(0.03, 0.03, 0.07, 0.06), **kw))
    f = tmp_path / "delam.mat"
    matio.save_mesh(f, a["nodes"], a["coords"], 5, delamEl=[a["den_under"], a["den_above"]],
                    den_under=[a["den_under"]], den_above=[a["den_above"]])
    b = E.prepare_flat_shell(E.shell_plate_40x40(mesh_file=str(f), **kw))
    for k in ("nodes", "coords", "coef", "elem_mat", "sensor_dofs", "ex_dofs", "st"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
