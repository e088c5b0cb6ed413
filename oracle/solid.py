"""3D-solid SEM oracle (TEST INFRASTRUCTURE): FP64 restatement of
``solid3D/main_spec3Dcuda_global_cpu_gpu3_wave_sens_opt_single_PCT_mat.m`` (the
"3D engine") and its helpers.

* ``lamina_q3d``          -- ``composite3D_transf.m:31-46`` orthotropic stiffness from
                             engineering constants (the reference gets them from
                             micromechanics, ``:20-30``; here they are given).
                             Pinned: isotropic limit gives lambda + 2 mu, lambda, mu.
* ``laminate_solid_smeared`` -- ``composite3D_transf.m:49-108`` ('smeared'), including the
                             reference's pairing D(5,5)=a55 with gamma_yz and D(6,6)=a44 with
                             gamma_xz (``:104-105``).  Pinned: for a single isotropic ply it
                             returns the isotropic D; cross-ply symmetric gives D14=D24=D34=0.
* ``SolidModel``          -- setup ``:69-95,129-145,394-604,666-811,853-888``: smeared D per
                             element, lumped mass ``mass3D_v2.m:17-64`` (w detJ rho), block-
                             diagonal ``Npx/Npy/Npz`` (``shape3D_prim_v2p.m``), 3x3 Jacobian
                             ``:480-488`` and its inverse/determinant (the reference's
                             ``inv_jacp``/``det_jacp`` are absent from its tree; supplied here
                             as the adjugate inverse -- the one function they can be),
                             ``Vandermonde_v2`` taken equal to ``Vandermonde.m``.
* ``SolidModel.internal_force`` -- loop body ``:927-1010``: gather, ``strains_spec_p.m:8-27``,
                             ``stresses_spec_p.m:6-11``, 27 transposed products ``:944-954``,
                             12-basket assembly.  Pinned: K symmetric PSD with the 6 rigid-body
                             modes as null space; patch test; brute force on tiny meshes.
* ``SolidModel.run``      -- ``:922-1144``: nn = nn_first..nft, F1 = Fa*ST(nn) (``:1018-1030``),
                             update ``:1032-1045``, ``displ`` = U after the update (``:1049-1051``).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .common import gll, vandermonde, shape3d_prim, nodes_maps, basket_assemble
from .shell import ply_z

SOLID_COEFS = ("D11", "D12", "D13", "D14", "D22", "D23", "D24", "D33", "D34", "D44", "D55", "D56", "D66", "rho")


def lamina_q3d(E11, E22, E33, G12, G13, G23, nu12, nu13, nu23):
    nu21 = E22 / E11 * nu12
    nu31 = E33 / E11 * nu13
    nu32 = E33 / E22 * nu23
    Delta = 1 - nu12 * nu21 - nu13 * nu31 - nu12 * nu23 * nu31 - nu13 * nu21 * nu32 - nu23 * nu32
    return dict(q11=E11 * (1 - nu23 * nu32) / Delta, q22=E22 * (1 - nu31 * nu13) / Delta,
                q33=E33 * (1 - nu12 * nu21) / Delta, q12=(nu21 + nu31 * nu23) * E11 / Delta,
                q13=(nu31 + nu21 * nu32) * E11 / Delta, q23=(nu32 + nu12 * nu31) * E22 / Delta,
                q44=G23, q55=G13, q66=G12)


def laminate_solid_smeared(lh, lalpha, rho, q11, q12, q13, q22, q23, q33, q44, q55, q66):
    h, h1, h2 = ply_z(lh)
    lay = len(h1)
    al = np.asarray(lalpha, dtype=np.float64) * np.pi / 180
    m, n = np.cos(al), np.sin(al)
    eps = np.finfo(float).eps
    m = np.where(np.abs(m) < eps, 0.0, m)
    n = np.where(np.abs(n) < eps, 0.0, n)
    b = lambda v: np.broadcast_to(np.asarray(v, float), (lay,))
    q11, q12, q13, q22, q23, q33, q44, q55, q66, rho = map(b, (q11, q12, q13, q22, q23, q33, q44, q55, q66, rho))
    s11 = q11 * m**4 + 2 * (q12 + 2 * q66) * m**2 * n**2 + q22 * n**4
    s12 = (q11 + q22 - 4 * q66) * m**2 * n**2 + q12 * (m**4 + n**4)
    s13 = q13 * m**2 + q23 * n**2
    s16 = (q11 - q12 - 2 * q66) * m**3 * n + (q12 - q22 + 2 * q66) * m * n**3
    s22 = q11 * n**4 + 2 * (q12 + 2 * q66) * m**2 * n**2 + q22 * m**4
    s23 = n**2 * q13 + m**2 * q23
    s33 = q33
    s26 = (q11 - q12 - 2 * q66) * n**3 * m + (q12 - q22 + 2 * q66) * n * m**3
    s36 = (q13 - q23) * m * n
    s44 = q44 * m**2 + q55 * n**2
    s45 = (q55 - q44) * m * n
    s55 = q55 * m**2 + q44 * n**2
    s66 = (q11 + q22 - 2 * q12 - 2 * q66) * m**2 * n**2 + q66 * (m**4 + n**4)
    f = (h2 - h1) / h
    avg = lambda s: float(np.sum(s * f))
    return dict(D11=avg(s11), D12=avg(s12), D13=avg(s13), D14=avg(s16), D22=avg(s22), D23=avg(s23),
                D24=avg(s26), D33=avg(s33), D34=avg(s36), D44=avg(s66), D55=avg(s55), D56=avg(s45),
                D66=avg(s44), rho=avg(rho))


def laminate_solid_layered(lalpha, rho, q11, q12, q13, q22, q23, q33, q44, q55, q66):
    """``composite3D_transf.m:109-119`` prop='layered': one row per ply (``J0 = rho``), the
    ply's rotated stiffness in the smeared row's layout and pairing (``:111-118``: D(5,5) = s55,
    D(5,6) = s45, D(6,6) = s44).  Returns a list of dicts (SOLID_COEFS keys), ply by ply."""
    lay = len(np.atleast_1d(lalpha))
    rows = []
    for i in range(lay):
        pick = lambda v: np.broadcast_to(np.asarray(v, float), (lay,))[i]  # noqa: E731
        # a one-ply laminate's smeared row is that ply's rotated stiffness (weight 1)
        r = laminate_solid_smeared([1.0], [np.atleast_1d(lalpha)[i]], pick(rho), pick(q11), pick(q12),
                                   pick(q13), pick(q22), pick(q23), pick(q33), pick(q44), pick(q55), pick(q66))
        rows.append(r)
    return rows


def inv_jacp(J11, J12, J13, J21, J22, J23, J31, J32, J33):
    det = J11 * (J22 * J33 - J23 * J32) - J12 * (J21 * J33 - J23 * J31) + J13 * (J21 * J32 - J22 * J31)
    a = 1.0 / det
    return (a * (J22 * J33 - J23 * J32), a * (J13 * J32 - J12 * J33), a * (J12 * J23 - J13 * J22),
            a * (J23 * J31 - J21 * J33), a * (J11 * J33 - J13 * J31), a * (J13 * J21 - J11 * J23),
            a * (J21 * J32 - J22 * J31), a * (J12 * J31 - J11 * J32), a * (J11 * J22 - J12 * J21)), det


class SolidModel:
    nf = 3

    def __init__(self, nodes, coords, nx, nz, coef, elem_mat=None, stiff_scale=None, rho_scale=None,
                 dv_field=(0.0, 0.0, 0.0)):
        nodes = np.asarray(nodes, dtype=np.int64)
        coords = np.asarray(coords, dtype=np.float64)
        self.nodes = nodes
        self.fen, self.nloc = nodes.shape
        self.nx, self.nz = nx, nz
        assert self.nloc == nx * nx * nz
        self.n_nodes = coords.shape[0]
        self.dof = 3 * self.n_nodes
        ksi, wx = gll(nx)
        dz, wz = gll(nz)
        Qx = vandermonde(ksi)
        Qz = vandermonde(dz)
        Nx, Ny, Nz = shape3d_prim(nx, nx, nz, Qx, Qx, Qz, ksi, ksi, dz)
        self.Npx = sp.block_diag([sp.csr_matrix(Nx)] * self.fen, format="csr")
        self.Npy = sp.block_diag([sp.csr_matrix(Ny)] * self.fen, format="csr")
        self.Npz = sp.block_diag([sp.csr_matrix(Nz)] * self.fen, format="csr")
        self.Npxt, self.Npyt, self.Npzt = (M.T.tocsr() for M in (self.Npx, self.Npy, self.Npz))
        I = nodes.reshape(-1)
        X, Y, Z = coords[I, 0], coords[I, 1], coords[I, 2]
        J11 = self.Npx @ X; J21 = self.Npx @ Y; J31 = self.Npx @ Z
        J12 = self.Npy @ X; J22 = self.Npy @ Y; J32 = self.Npy @ Z
        J13 = self.Npz @ X; J23 = self.Npz @ Y; J33 = self.Npz @ Z
        inv, detJ = inv_jacp(J11, J12, J13, J21, J22, J23, J31, J32, J33)
        (self.invJ11, self.invJ12, self.invJ13, self.invJ21, self.invJ22, self.invJ23,
         self.invJ31, self.invJ32, self.invJ33) = inv
        www = np.einsum("k,j,i->kji", wz, wx, wx).reshape(-1)
        self.WWWDetJ = np.tile(www, self.fen) * detJ
        coef = np.atleast_2d(np.asarray(coef, dtype=np.float64))
        elem_mat = np.zeros(self.fen, np.int64) if elem_mat is None else np.asarray(elem_mat, np.int64)
        ss = np.ones(self.fen) if stiff_scale is None else np.asarray(stiff_scale, np.float64)
        rs = np.ones(self.fen) if rho_scale is None else np.asarray(rho_scale, np.float64)
        per_el = coef[elem_mat]
        for k, nm in enumerate(SOLID_COEFS):
            sc = rs if nm == "rho" else ss
            setattr(self, nm, np.repeat(per_el[:, k] * sc, self.nloc))
        self.Iu, self.Iv, self.Iw = 3 * I, 3 * I + 1, 3 * I + 2
        self.I_G, self.I_L = nodes_maps(nodes, self.n_nodes)
        mq = np.repeat((self.WWWDetJ * self.rho)[:, None], 3, axis=1)
        self.mass = basket_assemble(self.I_G, self.I_L, mq, self.n_nodes, 3)
        self.dv = np.tile(np.asarray(dv_field, np.float64), self.n_nodes)

    def local_forces(self, U):
        Npx, Npy, Npz, Npxt, Npyt, Npzt = self.Npx, self.Npy, self.Npz, self.Npxt, self.Npyt, self.Npzt
        i11, i12, i13 = self.invJ11, self.invJ12, self.invJ13
        i21, i22, i23 = self.invJ21, self.invJ22, self.invJ23
        i31, i32, i33 = self.invJ31, self.invJ32, self.invJ33
        W = self.WWWDetJ
        UX, UY, UZ = U[self.Iu], U[self.Iv], U[self.Iw]
        # strains_spec_p.m:8-27
        Exx = (Npx @ UX) * i11 + (Npy @ UX) * i21 + (Npz @ UX) * i31
        Eyy = (Npx @ UY) * i12 + (Npy @ UY) * i22 + (Npz @ UY) * i32
        Ezz = (Npx @ UZ) * i13 + (Npy @ UZ) * i23 + (Npz @ UZ) * i33
        B41UX = (Npx @ UX) * i12 + (Npy @ UX) * i22 + (Npz @ UX) * i32
        B42UY = (Npx @ UY) * i11 + (Npy @ UY) * i21 + (Npz @ UY) * i31
        Exy = B41UX + B42UY
        B52UY = (Npx @ UY) * i13 + (Npy @ UY) * i23 + (Npz @ UY) * i33
        B53UZ = (Npx @ UZ) * i12 + (Npy @ UZ) * i22 + (Npz @ UZ) * i32
        Eyz = B52UY + B53UZ
        B61UX = (Npx @ UX) * i13 + (Npy @ UX) * i23 + (Npz @ UX) * i33
        B63UZ = (Npx @ UZ) * i11 + (Npy @ UZ) * i21 + (Npz @ UZ) * i31
        Exz = B61UX + B63UZ
        # stresses_spec_p.m:6-11
        Sxx = Exx * self.D11 + Eyy * self.D12 + Ezz * self.D13 + Exy * self.D14
        Syy = Exx * self.D12 + Eyy * self.D22 + Ezz * self.D23 + Exy * self.D24
        Szz = Exx * self.D13 + Eyy * self.D23 + Ezz * self.D33 + Exy * self.D34
        Sxy = Exx * self.D14 + Eyy * self.D24 + Ezz * self.D34 + Exy * self.D44
        Syz = Eyz * self.D55 + Exz * self.D56
        Sxz = Eyz * self.D56 + Exz * self.D66
        # :944-954
        fu = Npxt @ (Sxx * i11 * W) + Npyt @ (Sxx * i21 * W) + Npzt @ (Sxx * i31 * W)
        fu = fu + Npxt @ (Sxy * i12 * W) + Npyt @ (Sxy * i22 * W) + Npzt @ (Sxy * i32 * W)
        fu = fu + Npxt @ (Sxz * i13 * W) + Npyt @ (Sxz * i23 * W) + Npzt @ (Sxz * i33 * W)
        fv = Npxt @ (Syy * i12 * W) + Npyt @ (Syy * i22 * W) + Npzt @ (Syy * i32 * W)
        fv = fv + Npxt @ (Sxy * i11 * W) + Npyt @ (Sxy * i21 * W) + Npzt @ (Sxy * i31 * W)
        fv = fv + Npxt @ (Syz * i13 * W) + Npyt @ (Syz * i23 * W) + Npzt @ (Syz * i33 * W)
        fw = Npxt @ (Szz * i13 * W) + Npyt @ (Szz * i23 * W) + Npzt @ (Szz * i33 * W)
        fw = fw + Npxt @ (Syz * i12 * W) + Npyt @ (Syz * i22 * W) + Npzt @ (Syz * i32 * W)
        fw = fw + Npxt @ (Sxz * i11 * W) + Npyt @ (Sxz * i21 * W) + Npzt @ (Sxz * i31 * W)
        return np.stack([fu, fv, fw], axis=1)

    def internal_force(self, U):
        local = self.chunked if getattr(self, "chunked", None) is not None else self.local_forces
        return basket_assemble(self.I_G, self.I_L, local(U), self.n_nodes, 3)

    def run(self, dt, st, f_ext, nn_first=2, nn_last=None, outputs=(), out_quantity="disp_post",
            frame_no=(), frame_quantity="disp_post", out_weights=None):
        """Loop ``:922-1144``: F = -F + mg2 U + F1 - mg uold + a1 cg uold; unew = mg0 F;
        traces of U after the update (``:1049-1051``) unless ``out_quantity`` says otherwise.
        ``out_weights`` [n_out, dof] (instead of ``outputs``): traces W @ u after the update,
        the PZT sensors' voltage functionals (``:1056-1067`` as a linear functional)."""
        nft = len(st)
        nn_last = nft if nn_last is None else nn_last
        a0 = 1.0 / dt**2
        a1 = 1.0 / (2 * dt)
        mg_raw = self.mass
        cg = self.dv * mg_raw
        mg2 = 2 * a0 * mg_raw
        mg = a0 * mg_raw
        mg0 = 1.0 / (mg + a1 * cg)
        U = np.zeros(self.dof)
        uold = np.zeros(self.dof)
        outputs = np.asarray(outputs, dtype=np.int64)
        n_out = outputs.size if out_weights is None else np.shape(out_weights)[0]
        traces = np.zeros((nft, n_out))
        frame_no = list(frame_no)
        frames = np.zeros((len(frame_no), self.dof))
        c = 0
        for nn in range(nn_first, nn_last + 1):
            F = self.internal_force(U)
            F = -F + mg2 * U + f_ext * st[nn - 1] - mg * uold + a1 * cg * uold
            unew = mg0 * F
            if out_weights is not None:
                traces[nn - 1] = out_weights @ unew
            elif outputs.size:
                if out_quantity == "disp":
                    traces[nn - 1] = U[outputs]
                elif out_quantity == "vel":
                    traces[nn - 1] = a1 * (-uold[outputs] + unew[outputs])
                else:
                    traces[nn - 1] = unew[outputs]
            if c < len(frame_no) and nn == frame_no[c]:
                frames[c] = unew if frame_quantity == "disp_post" else a1 * (-uold + unew)
                c += 1
            uold = U
            U = unew
        return dict(traces=traces, frames=frames, U=U, uold=uold)
