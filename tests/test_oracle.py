"""Pins of the CPU oracle (no GPU): every check here is independent of the oracle's own
formulas -- golden data from the reference tree, closed forms, invariants and brute
force -- so that a dropped term, wrong sign, index or transposition fails one of them."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import common as oc, mesh as om, shell as osh, solid as oso

GOLD = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ A1-A4: GLL operators
@pytest.mark.parametrize("n", list(range(2, 11)))
def test_gll_matches_reference_table(n):
    """The oracle's nodes and weights are the fixture extracted from gll.m bit for bit; the
    table itself is pinned by the closed form (roots of P'_{n-1}, w = 2/(n(n-1)P_{n-1}^2))."""
    t = json.loads((GOLD / "gll_table.json").read_text())[str(n)]
    xo, wo = oc.gll(n)
    assert np.array_equal(xo, t["ksi"]) and np.array_equal(wo, t["wi"])
    assert not np.signbit(xo).any() or (np.signbit(xo) == (xo < 0)).all()  # no -0 in the middle
    x, w = oc.gll_closed_form(n)
    assert np.allclose(x, t["ksi"], rtol=0, atol=2e-15)
    # the reference's literal weights for n = 9, 10 carry truncations of 3.7e-15 and 1.0e-14
    # (their sums miss 2 by 6.7e-15 and 2.0e-14); the closed form sums to 2 within 1e-14
    assert np.allclose(w, t["wi"], rtol=0, atol=1e-15 if n <= 8 else 1.2e-14)
    assert abs(w.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7, 8])
def test_derivative_matrix_exact_on_polynomials(n):
    x, _ = oc.gll(n)
    Q = oc.vandermonde(x)
    S, Sp = oc.shape1d(Q, x)
    assert np.allclose(S, np.eye(n), atol=1e-12)  # Lagrange property L_j(x_i) = delta_ij
    for k in range(n):  # D exact on polynomials of degree <= n-1
        d = Sp @ x ** k
        ref = k * x ** (k - 1) if k else np.zeros(n)
        assert np.allclose(d, ref, atol=1e-11 * max(1, k * k))
    # closed form D_ij = P_N(x_i) / (P_N(x_j)(x_i - x_j)), D_00 = -N(N+1)/4
    N = n - 1
    P = np.polynomial.legendre.legval(x, [0] * N + [1])
    D = np.array([[P[i] / (P[j] * (x[i] - x[j])) if i != j else 0.0 for j in range(n)] for i in range(n)])
    D[0, 0], D[-1, -1] = -N * (N + 1) / 4, N * (N + 1) / 4
    assert np.allclose(Sp, D, atol=1e-11)


def test_shape2d_kronecker_structure():
    n = 4
    x, _ = oc.gll(n)
    Q = oc.vandermonde(x)
    Nx, Ny = oc.shape2d_prim(n, n, Q, Q, x, x)
    X, Y = np.meshgrid(x, x)  # point c = k2*n + k1: x varies fastest
    f = (X ** 2 * Y + 3 * Y ** 3).reshape(-1)
    assert np.allclose(Nx @ f, (2 * X * Y).reshape(-1), atol=1e-11)
    assert np.allclose(Ny @ f, (X ** 2 + 9 * Y ** 2).reshape(-1), atol=1e-11)


def test_hanning_signal_properties():
    """Properties a 5-cycle Hann-windowed burst has whatever its formula's form
    (Hanning_signal.m:17-23): support [t1, t1 + 1/f1]; the carrier at f2 (DFT peak, 9 sign
    changes inside the burst); antisymmetry about the burst centre (whole cycles under a
    symmetric window); a cubic onset (window ~ t^2, carrier ~ t); peak amplitude below 1 and
    energy sum st^2 dt = 3T/16 (mean of the Hann window squared 3/8, of sin^2 1/2).  A swapped
    f1/f2, a raised-cosine window of the wrong phase, a missing 1/2, a window over the wrong
    span or a one-sample time shift each fail one of them."""
    dt, nft, f2 = 1e-7, 4000, 50e3
    f1 = f2 / 5
    T = 1 / f1
    t, st = oc.hanning_signal(dt, nft, f1, f2, 0.0)
    n_T = int(round(T / dt))  # 1000 samples in the burst
    assert np.allclose(t, np.arange(nft) * dt)
    assert np.all(st[n_T + 1:] == 0.0) and st[0] == 0.0
    burst = st[:n_T + 1]
    # antisymmetry about the centre: st(T - tau) = -st(tau)
    assert np.allclose(burst, -burst[::-1], atol=1e-12)
    # carrier: sign changes strictly inside the burst and the DFT peak
    s_in = burst[1:-1]
    changes = np.count_nonzero(np.sign(s_in[1:]) * np.sign(s_in[:-1]) < 0)
    assert changes == 9
    spec = np.abs(np.fft.rfft(st, n=1 << 16))
    freqs = np.fft.rfftfreq(1 << 16, dt)
    assert abs(freqs[np.argmax(spec)] - f2) < 0.05 * f2
    # onset: |st(t)| <= C t^3 with C = pi^2 f1^2 * 2 pi f2
    k = np.arange(1, 20)
    assert np.all(np.abs(st[k]) <= 1.01 * (np.pi * f1) ** 2 * 2 * np.pi * f2 * (k * dt) ** 3)
    # amplitude and energy
    assert 0.9 < np.abs(burst).max() <= 1.0
    assert abs((st ** 2).sum() * dt / (3 * T / 16) - 1) < 1e-3


# ------------------------------------------------------------------ A6: laminate
def test_lamina_isotropic_closed_form():
    E, nu = 70e9, 0.33
    q = osh.lamina_q(E, E, E / (2 * (1 + nu)), E / (2 * (1 + nu)), E / (2 * (1 + nu)), nu)
    assert np.isclose(q["q11"], E / (1 - nu ** 2)) and np.isclose(q["q12"], nu * E / (1 - nu ** 2))


def test_laminate_known_answers():
    E11, E22, G12, G23, nu12, rho = 130e9, 10e9, 5e9, 3.5e9, 0.3, 1570
    Q = osh.lamina_q(E11, E22, G12, G12, G23, nu12)
    lh = [0.25e-3] * 8
    h = 2e-3
    c = osh.laminate_shell(lh, [0, 90, 0, 90, 90, 0, 90, 0], rho, **Q)
    Amax = c["a11"]
    for k in ("b11", "b12", "b16", "b22", "b26", "b66"):
        assert abs(c[k]) < 1e-12 * Amax * h
    assert abs(c["a16"]) < 1e-12 * Amax and abs(c["a26"]) < 1e-12 * Amax
    assert np.isclose(c["m11"], rho * h) and np.isclose(c["I11"], rho * h ** 3 / 12)
    # cross-ply, half of each orientation: A11 = A22 = (Q11+Q22)/2 h
    assert np.isclose(c["a11"], (Q["q11"] + Q["q22"]) / 2 * h)
    # homogeneous isotropic plate: D = E h^3 / 12 / (1-nu^2), shear = 5/6 G h
    E, nu = 70e9, 0.3
    G = E / (2 * (1 + nu))
    Qi = osh.lamina_q(E, E, G, G, G, nu)
    ci = osh.laminate_shell([h / 4] * 4, [0, 0, 0, 0], 2700, **Qi)
    assert np.isclose(ci["d11"], E * h ** 3 / 12 / (1 - nu ** 2))
    assert np.isclose(ci["a44"], 5 / 6 * G * h) and np.isclose(ci["a55"], 5 / 6 * G * h)
    # a symmetric angle-ply keeps B = 0 but has D16 != 0
    ca = osh.laminate_shell(lh, [45, -45, 45, -45, -45, 45, -45, 45], rho, **Q)
    assert abs(ca["b16"]) < 1e-12 * ca["a11"] * h and abs(ca["d16"]) > 1e-3 * ca["d11"]


def test_solid_smeared_isotropic_and_quirk():
    E, nu = 70e9, 0.3
    G = E / (2 * (1 + nu))
    Q = oso.lamina_q3d(E, E, E, G, G, G, nu, nu, nu)
    lam = oso.laminate_solid_smeared([1e-3], [0], 2700, **Q)
    lmb = E * nu / ((1 + nu) * (1 - 2 * nu))
    assert np.isclose(lam["D11"], lmb + 2 * G) and np.isclose(lam["D12"], lmb) and np.isclose(lam["D44"], G)
    # 0-degree orthotropic ply: D55 carries q55 (= G13) and D66 carries q44 (= G23), the
    # reference's pairing (composite3D_transf.m:104-105)
    Q = oso.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 0.4)
    lam = oso.laminate_solid_smeared([1e-3], [0], 1570, **Q)
    assert np.isclose(lam["D55"], 5e9) and np.isclose(lam["D66"], 3.5e9) and lam["D56"] == 0


# ------------------------------------------------------------------ A7-A8: index maps
def test_baskets_invariants_and_hand_case():
    q, p = om.rect_quad_mesh(2.0, 1.0, 2, 1)
    nodes, coords = om.quad2spectral(q, p, 2)  # 2 elements of 3x3 points, 15 nodes
    IG, IL = oc.nodes_maps(nodes, coords.shape[0])
    assert IG.shape == (2, 12) and (IG >= 0).sum() == 18  # 18 points: 6 empty slots
    # hand-worked: node 0 (single occurrence) -> basket 0 row 0; the shared corner node 1
    # is the first node with 2 occurrences -> baskets 1 and 2 (lowest fill, lowest index)
    assert IG[0, 0] == 0 and IL[0, 0] == 0
    assert IG[0, 1] == 1 and IG[0, 2] == 1 and sorted([IL[0, 1], IL[0, 2]]) == [2, 9]
    for b in range(12):
        col = IG[:, b][IG[:, b] >= 0]
        assert len(set(col)) == len(col)  # unique nodes per basket: race-free scatter
    flat = nodes.reshape(-1)
    pairs = {(int(IG[r, b]), int(IL[r, b])) for r in range(IG.shape[0]) for b in range(12) if IG[r, b] >= 0}
    assert pairs == {(int(flat[q_]), q_) for q_ in range(flat.size)}  # bijection


def test_quad2spectral_structured_grid_is_exact():
    q, p = om.rect_quad_mesh(0.5, 0.25, 4, 2)
    nodes, coords = om.quad2spectral(q, p, 4)
    x, _ = oc.gll(5)
    for e in range(nodes.shape[0]):
        c = coords[nodes[e]].reshape(5, 5, 3)
        x0, y0 = c[0, 0, 0], c[0, 0, 1]
        X = x0 + (x + 1) / 2 * 0.125
        Y = y0 + (x + 1) / 2 * 0.125
        assert np.allclose(c[:, :, 0], X[None, :], atol=1e-15)
        assert np.allclose(c[:, :, 1], Y[:, None], atol=1e-15)
    assert len(np.unique(nodes)) == coords.shape[0] == (4 * 4 + 1) * (2 * 4 + 1)


def test_quad2spectral_gmsh_fixture():
    m = json.loads((GOLD / "m1_rand_2.json").read_text())
    pos = np.array(m["POS"])
    quads = om.change_turn_quad(np.array(m["QUADS"])[:, :4] - 1, pos)
    nodes, coords = om.quad2spectral(quads, pos, 5)
    assert nodes.shape == (16, 36)
    assert len(np.unique(nodes)) == coords.shape[0]  # every node used
    # conformity: each shared edge has its 4 interior nodes shared by exactly two elements
    cnt = np.bincount(nodes.reshape(-1))
    ed = lambda e: set(nodes[e].reshape(6, 6)[0, 1:5]) | set(nodes[e].reshape(6, 6)[5, 1:5]) | \
        set(nodes[e].reshape(6, 6)[1:5, 0]) | set(nodes[e].reshape(6, 6)[1:5, 5])
    shared = [k for e in range(16) for k in ed(e) if cnt[k] == 2]
    assert len(shared) > 0 and all(cnt[k] in (1, 2) for e in range(16) for k in ed(e))
    # total area through the GLL quadrature = area of the polygon of the bilinear quads
    x, w = oc.gll(6)
    model = osh.ShellModel(nodes, coords, 5, np.ones((1, 23)))
    area = model.WWDetJ.sum()
    pq = pos[quads]
    quad_area = 0.5 * np.abs((pq[:, 2, 0] - pq[:, 0, 0]) * (pq[:, 3, 1] - pq[:, 1, 1])
                             - (pq[:, 3, 0] - pq[:, 1, 0]) * (pq[:, 2, 1] - pq[:, 0, 1])).sum()
    assert np.isclose(area, quad_area, rtol=1e-6)
    assert (model.detJ > 0).all()


# ------------------------------------------------------------------ A10-A14: operator
def _shell(nex=2, ney=2, order=3, lx=0.024, ly=0.02, perturb=0.0, angles=(0, 90, 0, 90, 90, 0, 90, 0)):
    q, p = om.rect_quad_mesh(lx, ly, nex, ney)
    if perturb:
        rng = np.random.default_rng(0)
        inner = (p[:, 0] > 0) & (p[:, 0] < lx) & (p[:, 1] > 0) & (p[:, 1] < ly)
        p = p.copy()
        p[inner, :2] += perturb * rng.uniform(-1, 1, (inner.sum(), 2)) * lx / nex
    n, c = om.quad2spectral(q, p, order)
    Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
    lam = osh.laminate_shell([0.25e-3] * len(angles), list(angles), 1570, **Q)
    return osh.ShellModel(n, c, order, np.array([[lam[k] for k in osh.SHELL_COEFS]])), c


@pytest.mark.parametrize("perturb,angles", [(0.0, (0, 90, 0, 90, 90, 0, 90, 0)),
                                            (0.15, (30, -60, 45, 0))])
def test_shell_K_symmetric_psd_rigid_modes(perturb, angles):
    m, c = _shell(perturb=perturb, angles=angles)
    K = np.stack([m.internal_force(e) for e in np.eye(m.dof)], axis=1)
    assert np.abs(K - K.T).max() < 1e-12 * np.abs(K).max()
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert w[0] > -1e-12 * w[-1]
    assert np.sum(np.abs(w) < 1e-12 * w[-1]) == 6
    x, y = c[:, 0], c[:, 1]
    z = np.zeros_like(x)
    o = np.ones_like(x)
    modes = [(o, z, z, z, z), (z, z, o, z, z), (z, z, z, z, o), (-y, z, x, z, z),
             (z, -o, z, z, x), (z, z, z, -o, y)]  # fields u, phix, v, phiy, w ('+' shear sign)
    for mo in modes:
        u = np.stack(mo, axis=1).reshape(-1)
        assert np.abs(K @ u).max() < 1e-10 * np.abs(K).max() * np.abs(u).max()


def test_shell_constant_strain_patch():
    m, c = _shell(nex=3, ney=3, order=4, perturb=0.15)
    x, y = c[:, 0], c[:, 1]
    z = np.zeros_like(x)
    u = np.stack([1e-4 * x + 2e-5 * y, z, -3e-5 * y + 1e-5 * x, z, z], axis=1).reshape(-1)
    F = m.internal_force(u).reshape(-1, 5)
    inner = (x > 1e-9) & (x < 0.024 - 1e-9) & (y > 1e-9) & (y < 0.02 - 1e-9)
    scale = np.abs(F).max()
    assert scale > 0
    assert np.abs(F[inner]).max() < 1e-9 * scale


def test_solid_K_symmetric_psd_rigid_modes():
    q, p = om.rect_quad_mesh(0.02, 0.02, 2, 1)
    n2, c2 = om.quad2spectral(q, p, 2)
    n3, c3 = om.extrude_mesh(n2, c2, 3, 3, 3, [1e-3, 1e-3], 2)
    Q = oso.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
    lam = oso.laminate_solid_smeared([0.25e-3] * 8, [0, 45, -45, 90, 90, -45, 45, 0], 1570, **Q)
    m = oso.SolidModel(n3, c3, 3, 3, [lam[k] for k in oso.SOLID_COEFS])
    K = np.stack([m.internal_force(e) for e in np.eye(m.dof)], axis=1)
    assert np.abs(K - K.T).max() < 1e-12 * np.abs(K).max()
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert np.sum(np.abs(w) < 1e-12 * w[-1]) == 6 and w[0] > -1e-12 * w[-1]
    x, y, zc = c3[:, 0], c3[:, 1], c3[:, 2]
    z0 = np.zeros_like(x)
    for mo in [(x * 0 + 1, z0, z0), (-y, x, z0), (zc, z0, -x)]:
        u = np.stack(mo, axis=1).reshape(-1)
        assert np.abs(K @ u).max() < 1e-9 * np.abs(K).max() * np.abs(u).max()
    assert np.isclose(m.mass[0::3].sum(), 1570 * 0.02 * 0.02 * 2e-3)


# ------------------------------------------------------------------ A9, A15-A17: loop
def test_mass_total():
    m, _ = _shell(nex=3, ney=2, order=5)
    assert np.isclose(m.mass[0::5].sum(), 1570 * 2e-3 * 0.024 * 0.02, rtol=1e-13)
    assert np.isclose(m.mass[1::5].sum(), 1570 * 2e-3 ** 3 / 12 * 0.024 * 0.02, rtol=1e-13)


def test_update_closed_form_recurrence():
    m, _ = _shell(nex=1, ney=1, order=2)
    for nm in osh.SHELL_COEFS[:21]:
        setattr(m, nm, np.zeros_like(getattr(m, nm)))  # K = 0: free masses
    dt = 1e-7
    _, st = oc.hanning_signal(dt, 50, 2e5, 1e6, 0.0)
    f = np.zeros(m.dof)
    f[4] = 2.0
    r = m.run(dt, st, f, 1, 50, [4], "disp")
    u_prev, u, ref = 0.0, 0.0, []
    for k in range(50):
        ref.append(u)
        un = 2 * u - u_prev + dt ** 2 * f[4] * st[k] / m.mass[4]
        u_prev, u = u, un
    assert np.allclose(r["traces"][:, 0], ref, rtol=1e-12, atol=1e-30)


def test_discrete_energy_conserved_after_burst():
    m, _ = _shell(nex=2, ney=2, order=4)
    dt = 2e-8
    nft = 600
    _, st = oc.hanning_signal(dt, nft, 1e6 / 5, 1e6, 0.0)  # burst over the first 250 steps
    f = np.zeros(m.dof)
    f[5 * 12 + 4] = 1.0
    K = lambda u: m.internal_force(u)
    E = []
    U, Uo = np.zeros(m.dof), np.zeros(m.dof)
    for nn in range(300, nft + 1, 50):
        r = m.run(dt, st, f, 1, nn, u0=None)
        Un, Uc = r["U"], r["uold"]  # U^{nn+1}, U^{nn}
        v = (Un - Uc) / dt
        E.append(0.5 * v @ (m.mass * v) + 0.5 * Un @ K(Uc))
    E = np.array(E)
    assert E[0] > 0
    assert np.abs(E - E[0]).max() < 1e-9 * E[0]


def test_reciprocity_point_force_response():
    m, c = _shell(nex=3, ney=3, order=3, perturb=0.1)
    dt = 5e-8
    _, st = oc.hanning_signal(dt, 200, 4e5 / 3, 4e5, 0.0)
    a, b = 5 * 7 + 4, 5 * 30 + 2
    fa = np.zeros(m.dof); fa[a] = 1.0
    fb = np.zeros(m.dof); fb[b] = 1.0
    ra = m.run(dt, st, fa, 1, 200, [b], "disp")
    rb = m.run(dt, st, fb, 1, 200, [a], "disp")
    assert np.abs(ra["traces"]).max() > 0
    assert np.allclose(ra["traces"], rb["traces"], rtol=1e-10, atol=1e-12 * np.abs(ra["traces"]).max())


@pytest.mark.parametrize("n", list(range(2, 11)))
def test_library_gll_is_the_reference_table(n):
    """The library's GLL rule (sem_gll, host code: no GPU) is gll.m's literal table bit for
    bit, as the oracle's is: the mesh front end's node coordinates, and the order in which
    unique(...,'rows') numbers edge nodes, follow from these values."""
    from semhip import pzt
    t = json.loads((GOLD / "gll_table.json").read_text())[str(n)]
    x, w = pzt.gll_nodes(n)
    assert np.array_equal(x, t["ksi"]) and np.array_equal(w, t["wi"])


def test_quad2spectral_numbering_on_a_symmetric_fan():
    """A ten-blade fan of kite quads around one node: mirror-image edge nodes share their x
    coordinate, so unique(...,'rows') orders them by y only if both x values are the same
    double.  The library numbers the nodes as the oracle does, bit for bit, at every order."""
    import semhip
    n_bl, r, rm = 10, 0.01, 0.013
    ang = 2 * np.pi * np.arange(n_bl) / n_bl
    v = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(n_bl)], axis=1)
    m = np.stack([rm * np.cos(ang + np.pi / n_bl), rm * np.sin(ang + np.pi / n_bl), np.zeros(n_bl)], axis=1)
    pos = np.concatenate([np.zeros((1, 3)), v, m])
    k = np.arange(n_bl)
    quads = np.stack([np.zeros(n_bl, int), 1 + k, 1 + n_bl + k, 1 + (k + 1) % n_bl], axis=1)
    for order in range(2, 10):
        n1, c1 = om.quad2spectral(quads, pos, order)
        n2, c2 = semhip.quad2spectral(quads, pos, order)
        assert np.array_equal(n1, n2), order
        assert np.array_equal(c1, c2), order
