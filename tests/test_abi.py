"""CPU checks of the C-ABI library: it loads, exports every symbol include/sem_hip.h
declares, and its host-only entry points (mesh front end, index maps, laminate
helpers) agree with the oracle.  No GPU compute is called here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import semhip
from semhip import _lib as L
from oracle import mesh as om, common as oc, shell as osh, solid as oso

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    txt = (ROOT / "include" / "sem_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(sem_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol(lib_path):
    assert lib_path.exists(), "build first: python ma-shm_amd/build.py"
    lib = ctypes.CDLL(str(lib_path))
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(L.SIGNATURES), set(syms) ^ set(L.SIGNATURES)


def test_version_string():
    assert b"gfx950" in L.load().sem_version()


@pytest.mark.parametrize("order", [2, 3, 4, 5, 6, 8])
def test_quad2spectral_structured_matches_oracle(order):
    q, p = semhip.rect_quads(0.3, 0.2, 7, 5)
    n, c = semhip.quad2spectral(q, p, order)
    q2, p2 = om.rect_quad_mesh(0.3, 0.2, 7, 5)
    n2, c2 = om.quad2spectral(q2, p2, order)
    assert np.array_equal(n, n2)
    assert np.abs(c - c2).max() < 1e-15


def test_baskets_match_oracle_bit_exact():
    q, p = semhip.rect_quads(0.5, 0.5, 9, 7)
    n, c = semhip.quad2spectral(q, p, 5)
    IG, IL = semhip.baskets(n, c.shape[0])
    IG2, IL2 = oc.nodes_maps(n, c.shape[0])
    assert np.array_equal(IG, IG2 + 1) and np.array_equal(IL, IL2 + 1)


def test_extrude_matches_oracle():
    q, p = semhip.rect_quads(0.05, 0.04, 3, 2)
    n2, c2 = semhip.quad2spectral(q, p, 4)
    n3, c3 = semhip.extrude(n2, c2, 5, 5, [1e-3, 1e-3])
    o3, oc3 = om.extrude_mesh(n2, c2, 5, 5, 5, [1e-3, 1e-3], 2)
    assert np.array_equal(n3, o3)
    assert np.abs(c3 - oc3).max() < 1e-18


def test_laminate_shell_matches_oracle():
    q = semhip.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    c = semhip.laminate_shell([0.25e-3] * 8, [0, 90, 0, 90, 90, 0, 90, 0], 1570, q)
    Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    lam = osh.laminate_shell([0.25e-3] * 8, [0, 90, 0, 90, 90, 0, 90, 0], 1570, **Q)
    ref = np.array([lam[k] for k in osh.SHELL_COEFS])
    scale = np.maximum(np.abs(ref), 1e-12 * np.abs(ref).max())
    assert np.all(np.abs(c - ref) <= 1e-13 * np.abs(ref).max() + 1e-14 * scale)


def test_laminate_solid_matches_oracle():
    q3 = semhip.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
    ang = [0, 45, -45, 90, 90, -45, 45, 0]
    c = semhip.laminate_solid_smeared([0.25e-3] * 8, ang, 1570, q3)
    Q = oso.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
    lam = oso.laminate_solid_smeared([0.25e-3] * 8, ang, 1570, **Q)
    ref = np.array([lam[k] for k in oso.SOLID_COEFS])
    assert np.abs(c - ref).max() <= 1e-12 * np.abs(ref).max()


def test_laminate_solid_layered_rows():
    """Per-ply rows (composite3D_transf.m:109-119): a 0-degree ply is its lamina (with the
    reference's D55 = q55, D66 = q44 pairing), a 90-degree ply swaps the in-plane axes, a
    45-degree ply has s11 = (q11 + q22 + 2 q12 + 4 q66) / 4 and s16 = (q11 - q22) / 4; the
    library equals the oracle, and the thickness-weighted average of the rows is the smeared
    row of the same laminate (the identity between the two modes)."""
    q3 = semhip.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
    q11, q12, q13, q22, q23, q33, q44, q55, q66 = q3
    ang = [0, 90, 45, -45, 30, 0, 90, 60]
    rho = [1570, 1600, 1550, 1570, 1500, 1570, 1600, 1580]
    lh = [0.25e-3, 0.3e-3, 0.2e-3, 0.25e-3, 0.25e-3, 0.1e-3, 0.25e-3, 0.4e-3]
    rows = semhip.laminate_solid_layered(ang, rho, q3)
    assert rows.shape == (8, 14)
    order = ["D11", "D12", "D13", "D14", "D22", "D23", "D24", "D33", "D34", "D44", "D55", "D56", "D66", "rho"]
    r0 = dict(zip(order, rows[0]))
    assert np.allclose([r0[k] for k in order], [q11, q12, q13, 0, q22, q23, 0, q33, 0, q66, q55, 0, q44, 1570],
                       rtol=1e-14, atol=1e-3)
    r90 = dict(zip(order, rows[1]))
    assert np.allclose([r90[k] for k in order], [q22, q12, q23, 0, q11, q13, 0, q33, 0, q66, q44, 0, q55, 1600],
                       rtol=1e-12, atol=1e-3)
    r45 = dict(zip(order, rows[2]))
    assert np.isclose(r45["D11"], (q11 + q22 + 2 * q12 + 4 * q66) / 4, rtol=1e-13)
    assert np.isclose(r45["D14"], (q11 - q22) / 4, rtol=1e-13)
    assert np.isclose(r45["D56"], (q55 - q44) / 2, rtol=1e-13)
    Q = oso.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
    ref = np.array([[r[k] for k in oso.SOLID_COEFS] for r in oso.laminate_solid_layered(ang, rho, **Q)])
    assert np.abs(rows - ref).max() <= 1e-12 * np.abs(ref).max()
    f = np.asarray(lh) / np.sum(lh)
    smeared = semhip.laminate_solid_smeared(lh, ang, rho, q3)
    assert np.allclose(f @ rows, smeared, rtol=1e-13, atol=1e-3)


def test_engine_layered_solid_setup():
    """SolidCase(prop="layered") gives element layer k (from the bottom, the extrusion's
    element order) the k-th ply's row; a mesh with fewer element layers than plies is refused."""
    from semhip import engine as E
    case = E.SolidCase(lx=0.03, ly=0.02, nex=3, ney=2, nlay_el=4, h=2e-3, order=3, order_z=3,
                       lh=(0.5e-3,) * 4, angles=(0, 45, -45, 90), prop="layered", nft=50)
    prep = E.prepare_solid(case)
    q3 = semhip.lamina_q3d(case.E11, case.E22, case.E22, case.G12, case.G13, case.G23, case.nu12, case.nu12,
                           case.E22 / (2 * case.G23) - 1.0)
    assert np.array_equal(prep["coef"], semhip.laminate_solid_layered(case.angles, case.rho, q3))
    assert np.array_equal(prep["elem_mat"], np.repeat(np.arange(4), 6))
    z = prep["coords"][:, 2][prep["nodes"]].mean(axis=1)
    assert np.all(np.diff(z.reshape(4, 6).mean(axis=1)) > 0)  # layers in rising z
    smeared = E.prepare_solid(E.SolidCase(lx=0.03, ly=0.02, nex=3, ney=2, nlay_el=2, order=3, order_z=3, nft=50))
    assert smeared["coef"].shape == (1, 14) and "elem_mat" not in smeared
    with pytest.raises(ValueError):
        E.prepare_solid(E.SolidCase(lx=0.03, ly=0.02, nex=3, ney=2, nlay_el=2, order=3, order_z=3,
                                    prop="layered", nft=50))


def test_gmsh_fixture_through_the_product_mesher():
    """The reference-held gmsh mesh (gmsh_out/m1_rand_2.m, 16 quads) through the library's
    change_turn_quad and quad2spectral, with every third quad given clockwise: the library
    renumbers exactly those (back to the fixture's order) and both steps equal the oracle bit
    for bit at orders 3 and 5."""
    import json
    m = json.loads((ROOT / "tests" / "golden" / "m1_rand_2.json").read_text())
    pos = np.array(m["POS"])
    quads = np.array(m["QUADS"])[:, :4] - 1
    rev = quads.copy()
    rev[::3] = rev[::3][:, [0, 3, 2, 1]]
    qp, nflip = semhip.change_turn_quad(rev, pos)
    qo = om.change_turn_quad(rev, pos)
    assert nflip == len(rev[::3]) and np.array_equal(qp, qo) and np.array_equal(qp, quads)
    assert semhip.change_turn_quad(quads, pos)[1] == 0
    for order in (3, 5):
        n_p, c_p = semhip.quad2spectral(qp, pos, order)
        n_o, c_o = om.quad2spectral(qo, pos, order)
        assert np.array_equal(n_p, n_o) and np.array_equal(c_p, c_o)
