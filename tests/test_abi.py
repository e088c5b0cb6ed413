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
