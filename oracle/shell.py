"""Flat-shell SEM oracle (TEST INFRASTRUCTURE): FP64 restatement of
``flat_shell_engine/main_flat_shell_force_excit_c.m`` and its helpers.

What each part follows and what pins it (checks live in ``tests/``):

* ``lamina_q``           -- ``composite_plate.m:23-28`` (engineering constants -> reduced Q).
                            Pinned: isotropic case gives E/(1-nu^2), nu E/(1-nu^2).
* ``laminate_shell``     -- ``composite_plate_c_delam.m:10-19,48-87`` (ply rotation and the
                            A/B/D/shear/mass integrals, ply 1 on top: ``force_excit_c.m:75-83``).
                            Pinned: B=0 for symmetric layups, A16=A26=0 for cross-ply,
                            m11=rho h, I11=rho h^3/12, homogeneous D = E h^3/(12(1-nu^2)),
                            shear factor 5/6 for a homogeneous plate.
* ``ShellModel``         -- setup of ``force_excit_c.m:231-546``: block-diagonal sparse
                            ``Npx/Npy`` (``:231-300``), Jacobians ``:328-346`` with
                            ``inv_jacp_2D.m``/``det_jacp_2D.m``, ``WWDetJ`` ``:406-416``,
                            gather indices ``:424-437``, 12 baskets ``:480``, lumped mass
                            ``:497-503``, damping and central-difference constants ``:517-546``.
                            Pinned: total mass = rho h area; affine geometry closed form.
* ``ShellModel.internal_force`` -- the loop body ``force_excit_c.m:596-668`` (gather, 10
                            derivative products, stress resultants with the reference's own
                            expression order, 20 transposed products, 12-basket assembly).
                            Pinned: K symmetric and PSD with the 6-dimensional rigid null
                            space on a free plate, constant-strain patch test, brute force on
                            tiny meshes.
* ``ShellModel.run``     -- the time loop ``force_excit_c.m:592-862`` (update ``:687-693``,
                            sensors ``:701-705``, frames/divergence guard ``:710-730``).
                            Pinned: closed-form undamped recurrence on one DOF, discrete
                            energy conservation after the burst, source/receiver reciprocity.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .common import gll, vandermonde, shape2d_prim, nodes_maps, basket_assemble

SHELL_COEFS = ("a11", "a12", "a16", "a22", "a26", "a66", "a44", "a45", "a55",
               "b11", "b12", "b16", "b22", "b26", "b66",
               "d11", "d12", "d16", "d22", "d26", "d66", "m11", "I11")


def lamina_q(E11, E22, G12, G13, G23, nu12):
    """``composite_plate.m:23-28`` reduced stiffness of one lamina."""
    nu21 = E22 / E11 * nu12
    den = 1 - nu12 * nu21
    return dict(q11=E11 / den, q22=E22 / den, q12=nu12 * E22 / den, q44=G23, q55=G13, q66=G12)


def ply_z(lh):
    """Ply bounds with ply 1 on top, ``force_excit_c.m:75-83``."""
    lh = np.asarray(lh, dtype=np.float64)
    lay = lh.size
    h = float(np.sum(lh))
    h2 = np.zeros(lay)
    h1 = np.zeros(lay)
    h2[0] = h / 2
    h1[0] = h / 2 - lh[0]
    for i in range(1, lay):
        h2[i] = h2[i - 1] - lh[i - 1]
        h1[i] = h1[i - 1] - lh[i]
    return h, h1, h2


def laminate_shell(lh, lalpha, rho, q11, q12, q22, q44, q55, q66, plies=None):
    """``composite_plate_c_delam.m:10-19,48-87`` for an undelaminated element; with
    ``plies`` (0-based ply indices) the sums run over those plies only, at their z in the
    whole laminate and with the whole thickness in the shear factor: the elements above a
    delamination after ply ``L`` (1-based, the reference's ``delamination_layer``) take
    plies ``range(L)`` (``:89-131``), those under it ``range(L, lay)`` (``:133-173``).

    ``rho``/``q*`` may be scalars or per-ply arrays.  Returns a dict of the 23
    per-element coefficients (``SHELL_COEFS``)."""
    h, h1, h2 = ply_z(lh)
    lay = len(h1)
    al = np.asarray(lalpha, dtype=np.float64) * np.pi / 180
    m, n = np.cos(al), np.sin(al)
    q11, q12, q22, q44, q55, q66 = (np.broadcast_to(np.asarray(v, float), (lay,)) for v in (q11, q12, q22, q44, q55, q66))
    rho = np.broadcast_to(np.asarray(rho, float), (lay,))
    s11 = q11 * m**4 + 2 * (q12 + 2 * q66) * m**2 * n**2 + q22 * n**4
    s12 = (q11 + q22 - 4 * q66) * m**2 * n**2 + q12 * (m**4 + n**4)
    s16 = (q11 - q12 - 2 * q66) * m**3 * n + (q12 - q22 + 2 * q66) * m * n**3
    s22 = q11 * n**4 + 2 * (q12 + 2 * q66) * m**2 * n**2 + q22 * m**4
    s26 = (q11 - q12 - 2 * q66) * n**3 * m + (q12 - q22 + 2 * q66) * n * m**3
    s44 = q44 * m**2 + q55 * n**2
    s45 = (q55 - q44) * m * n
    s55 = q55 * m**2 + q44 * n**2
    s66 = (q11 + q22 - 2 * q12) * m**2 * n**2 + q66 * (m**2 - n**2) ** 2
    c = {k: 0.0 for k in SHELL_COEFS}
    for i in (range(lay) if plies is None else plies):
        t1 = h2[i] - h1[i]
        t2 = (h2[i] ** 2 - h1[i] ** 2) / 2
        t3 = (h2[i] ** 3 - h1[i] ** 3) / 3
        ts = 5 / 4 * (h2[i] - h1[i] - 4 / 3 * (h2[i] ** 3 - h1[i] ** 3) / h**2)
        c["m11"] += rho[i] * t1
        c["I11"] += rho[i] * t3
        c["a44"] += s44[i] * ts
        c["a45"] += s45[i] * ts
        c["a55"] += s55[i] * ts
        for nm, s in (("11", s11), ("12", s12), ("16", s16), ("22", s22), ("26", s26), ("66", s66)):
            c["a" + nm] += s[i] * t1
            c["b" + nm] += s[i] * t2
            c["d" + nm] += s[i] * t3
    return c


class ShellModel:
    """Setup + loop of ``main_flat_shell_force_excit_c.m`` on a given spectral mesh.

    ``coef`` is a [n_mat x 23] table (``SHELL_COEFS`` order) with ``elem_mat`` giving each
    element's row (the reference fills its per-point arrays element by element,
    ``composite_plate_c_delam.m:41-87``).  ``stiff_scale``/``rho_scale`` are per-element
    factors on the stiffness / mass coefficients (damage proxy; 1 reproduces the reference).
    ``dv_field`` is the damping factor per field u, phix, v, phiy, w (``:517-525``).
    """

    nf = 5

    def __init__(self, nodes, coords, order, coef, elem_mat=None, stiff_scale=None,
                 rho_scale=None, dv_field=(0.0, 0.0, 0.0, 0.0, 0.0)):
        nodes = np.asarray(nodes, dtype=np.int64)
        coords = np.asarray(coords, dtype=np.float64)
        self.nodes = nodes
        self.fen, self.nloc = nodes.shape
        self.n = order + 1
        assert self.nloc == self.n ** 2
        self.n_nodes = coords.shape[0]
        self.dof = 5 * self.n_nodes
        n = self.n
        ksi, wx = gll(n)
        Q = vandermonde(ksi)
        Nx, Ny = shape2d_prim(n, n, Q, Q, ksi, ksi)
        self.Nx_loc, self.Ny_loc = Nx, Ny
        blk = [sp.csr_matrix(Nx)] * self.fen
        self.Npx = sp.block_diag(blk, format="csr")
        self.Npy = sp.block_diag([sp.csr_matrix(Ny)] * self.fen, format="csr")
        self.Npxt = self.Npx.T.tocsr()
        self.Npyt = self.Npy.T.tocsr()
        I = nodes.reshape(-1)
        X = coords[I, 0]
        Y = coords[I, 1]
        J11 = self.Npx @ X
        J21 = self.Npx @ Y
        J12 = self.Npy @ X
        J22 = self.Npy @ Y
        a = 1.0 / (J11 * J22 - J12 * J21)
        self.invJ11 = a * J22
        self.invJ12 = a * (-J12)
        self.invJ21 = a * (-J21)
        self.invJ22 = a * J11
        detJ = J11 * J22 - J12 * J21
        self.detJ = detJ
        ww = np.outer(wx, wx).reshape(-1)  # ww(c) = wx(kx)*wy(ky), c = ky*n + kx
        self.WWDetJ = np.tile(ww, self.fen) * detJ
        # per-point coefficient arrays
        coef = np.atleast_2d(np.asarray(coef, dtype=np.float64))
        elem_mat = np.zeros(self.fen, np.int64) if elem_mat is None else np.asarray(elem_mat, np.int64)
        ss = np.ones(self.fen) if stiff_scale is None else np.asarray(stiff_scale, np.float64)
        rs = np.ones(self.fen) if rho_scale is None else np.asarray(rho_scale, np.float64)
        per_el = coef[elem_mat]  # [fen, 23]
        for k, nm in enumerate(SHELL_COEFS):
            scale = rs if nm in ("m11", "I11") else ss
            setattr(self, nm, np.repeat(per_el[:, k] * scale, self.nloc))
        self.Iu = 5 * I
        self.Ifix = 5 * I + 1
        self.Iv = 5 * I + 2
        self.Ifiy = 5 * I + 3
        self.Iw = 5 * I + 4
        self.I_G, self.I_L = nodes_maps(nodes, self.n_nodes)
        mgL = np.zeros((self.fen * self.nloc, 5))
        mgL[:, 0] = self.WWDetJ * self.m11
        mgL[:, 1] = self.WWDetJ * self.I11
        mgL[:, 2] = self.WWDetJ * self.m11
        mgL[:, 3] = self.WWDetJ * self.I11
        mgL[:, 4] = self.WWDetJ * self.m11
        self.mass = basket_assemble(self.I_G, self.I_L, mgL, self.n_nodes, 5)
        self.dv = np.tile(np.asarray(dv_field, np.float64), self.n_nodes)

    # ------------------------------------------------------------------ operator
    def local_forces(self, U):
        """Element-local forces [L x 5] (fu, mx, fv, my, fw), ``force_excit_c.m:596-630``."""
        Npx, Npy, Npxt, Npyt = self.Npx, self.Npy, self.Npxt, self.Npyt
        invJ11, invJ12, invJ21, invJ22, W = self.invJ11, self.invJ12, self.invJ21, self.invJ22, self.WWDetJ
        UX = U[self.Iu]; UY = U[self.Iv]; UZ = U[self.Iw]; FIX = U[self.Ifix]; FIY = U[self.Ifiy]
        NpxUX = Npx @ UX; NpxUY = Npx @ UY; NpxFIX = Npx @ FIX; NpxFIY = Npx @ FIY
        NpyUX = Npy @ UX; NpyUY = Npy @ UY; NpyFIX = Npy @ FIX; NpyFIY = Npy @ FIY
        NpxUZ = Npx @ UZ; NpyUZ = Npy @ UZ
        a11, a12, a16, a22, a26, a66 = self.a11, self.a12, self.a16, self.a22, self.a26, self.a66
        b11, b12, b16, b22, b26, b66 = self.b11, self.b12, self.b16, self.b22, self.b26, self.b66
        d11, d12, d16, d22, d26, d66 = self.d11, self.d12, self.d16, self.d22, self.d26, self.d66
        a44, a45, a55 = self.a44, self.a45, self.a55
        ux_x = NpxUX * invJ11 + NpyUX * invJ21
        fx_x = NpxFIX * invJ11 + NpyFIX * invJ21
        uy_y = NpxUY * invJ12 + NpyUY * invJ22
        fy_y = NpxFIY * invJ12 + NpyFIY * invJ22
        ux_y = NpxUX * invJ12 + NpyUX * invJ22
        fx_y = NpxFIX * invJ12 + NpyFIX * invJ22
        uy_x = NpxUY * invJ11 + NpyUY * invJ21
        fy_x = NpxFIY * invJ11 + NpyFIY * invJ21
        # :615-622, term order as coded
        Sxx = ux_x * a11 + fx_x * b11 + uy_y * a12 + fy_y * b12 + ux_y * a16 + fx_y * b16 + uy_x * a16 + fy_x * b16
        Sxxz = ux_x * b11 + fx_x * d11 + uy_y * b12 + fy_y * d12 + ux_y * b16 + fx_y * d16 + uy_x * b16 + fy_x * d16
        Sxy = ux_x * a16 + fx_x * b16 + uy_y * a26 + fy_y * b26 + ux_y * a66 + fx_y * b66 + uy_x * a66 + fy_x * b66
        Sxyz = ux_x * b16 + fx_x * d16 + uy_y * b26 + fy_y * d26 + ux_y * b66 + fx_y * d66 + uy_x * b66 + fy_x * d66
        Syy = ux_x * a12 + fx_x * b12 + uy_y * a22 + fy_y * b22 + ux_y * a26 + fx_y * b26 + uy_x * a26 + fy_x * b26
        Syyz = ux_x * b12 + fx_x * d12 + uy_y * b22 + fy_y * d22 + ux_y * b26 + fx_y * d26 + uy_x * b26 + fy_x * d26
        Syz = (FIY + NpxUZ * invJ12 + NpyUZ * invJ22) * a44 + (FIX + NpxUZ * invJ11 + NpyUZ * invJ21) * a45
        Sxz = (FIY + NpxUZ * invJ12 + NpyUZ * invJ22) * a45 + (FIX + NpxUZ * invJ11 + NpyUZ * invJ21) * a55
        # :624-630
        fu = Npxt @ (Sxx * invJ11 * W) + Npyt @ (Sxx * invJ21 * W) + Npxt @ (Sxy * invJ12 * W) + Npyt @ (Sxy * invJ22 * W)
        fv = Npxt @ (Syy * invJ12 * W) + Npyt @ (Syy * invJ22 * W) + Npxt @ (Sxy * invJ11 * W) + Npyt @ (Sxy * invJ21 * W)
        mx = (Npxt @ (Sxxz * invJ11 * W) + Npyt @ (Sxxz * invJ21 * W) + Npxt @ (Sxyz * invJ12 * W)
              + Npyt @ (Sxyz * invJ22 * W) + Sxz * W)
        my = (Npxt @ (Syyz * invJ12 * W) + Npyt @ (Syyz * invJ22 * W) + Npxt @ (Sxyz * invJ11 * W)
              + Npyt @ (Sxyz * invJ21 * W) + Syz * W)
        fw = Npxt @ (Syz * invJ12 * W) + Npyt @ (Syz * invJ22 * W) + Npxt @ (Sxz * invJ11 * W) + Npyt @ (Sxz * invJ21 * W)
        return np.stack([fu, mx, fv, my, fw], axis=1)

    def internal_force(self, U):
        """F_int = K U assembled over the 12 baskets (``force_excit_c.m:660-668``)."""
        local = self.chunked if getattr(self, "chunked", None) is not None else self.local_forces
        return basket_assemble(self.I_G, self.I_L, local(U), self.n_nodes, 5)

    # ------------------------------------------------------------------ time loop
    def run(self, dt, st, f_ext, nn_first=1, nn_last=None, outputs=(), out_quantity="disp",
            frame_no=(), frame_quantity="velocity", h_guard=None, u0=None, uold0=None):
        """Central-difference loop, ``force_excit_c.m:592-862``.

        ``st`` is the signal (1-based step nn reads st[nn-1]); ``f_ext`` the [dof]
        excitation vector (``Fi``).  Steps nn = nn_first..nn_last (force_excit
        convention: 1..nft).  ``outputs`` are 0-based DOFs; ``out_quantity``:
        'disp' = U before the update (``:701-705``), 'vel' = a1 (unew - uold),
        'disp_post' = unew.  Returns dict(traces[nft x n_out], frames[k x dof],
        t_frames, status, nn_done)."""
        nft = len(st)
        nn_last = nft if nn_last is None else nn_last
        a0 = 1.0 / dt**2
        a1 = 1.0 / (2 * dt)
        a2 = 2 * a0
        mg_raw = self.mass
        cg = self.dv * mg_raw
        mg2 = a2 * mg_raw
        mg = a0 * mg_raw
        mg0 = 1.0 / (mg + a1 * cg)
        U = np.zeros(self.dof) if u0 is None else np.array(u0, dtype=np.float64)
        uold = np.zeros(self.dof) if uold0 is None else np.array(uold0, dtype=np.float64)
        outputs = np.asarray(outputs, dtype=np.int64)
        traces = np.zeros((nft, outputs.size))
        frame_no = list(frame_no)
        frames = np.zeros((len(frame_no), self.dof))
        t_frames = np.zeros(len(frame_no))
        c = 0
        status = "ok"
        nn_done = nn_first - 1
        for nn in range(nn_first, nn_last + 1):
            F = self.internal_force(U)
            F = -F + mg2 * U + f_ext * st[nn - 1] - mg * uold + a1 * cg * uold
            unew = mg0 * F
            if outputs.size:
                if out_quantity == "disp":
                    traces[nn - 1] = U[outputs]
                elif out_quantity == "vel":
                    traces[nn - 1] = a1 * (-uold[outputs] + unew[outputs])
                else:
                    traces[nn - 1] = unew[outputs]
            if c < len(frame_no) and nn == frame_no[c]:
                t_frames[c] = (nn - 1) * dt
                if frame_quantity == "velocity":
                    frames[c] = a1 * (-uold + unew)
                elif frame_quantity == "displacement":
                    frames[c] = U
                else:
                    frames[c] = a0 * (uold - 2 * U + unew)
                UZ = U[self.Iw]
                # MATLAB's max omits NaN (``force_excit_c.m:712``)
                fin = UZ[~np.isnan(UZ)]
                umax = fin.max() if fin.size else np.nan
                if h_guard is not None and (umax > 10 * h_guard or np.isnan(unew[2])):
                    status = "diverged"
                    nn_done = nn
                    break
                c += 1
            uold = U
            U = unew
            nn_done = nn
        return dict(traces=traces, frames=frames, t_frames=t_frames, status=status,
                    nn_done=nn_done, U=U, uold=uold)


def nearest_node(coords, xy):
    """``force_excit_c.m:63-72``: ``min`` of the Euclidean distance, first index wins."""
    d = np.sqrt((coords[:, 0] - xy[0]) ** 2 + (coords[:, 1] - xy[1]) ** 2)
    return int(np.argmin(d))
