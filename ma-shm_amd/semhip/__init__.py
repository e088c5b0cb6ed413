"""semhip -- MI355X spectral-element time stepping (host side of libsemhip.so).

The compute path is the HIP library ``libsemhip.so`` (C-ABI ``include/sem_hip.h``);
this package is its Python host mirror: mesh/material helpers and the ``Solver``
context object.  ``engine`` mirrors the reference's MATLAB engine entry points.
There is no CPU fallback: every call goes through the HIP library and raises if it
is missing or fails.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import (SemError, SEM_SHELL5, SEM_SOLID3, SEM_FP64, SEM_FP32_SPLIT, SEM_Q_DISP,
                   SEM_Q_VEL, SEM_Q_DISP_POST, SEM_Q_ACC, SEM_E_DIVERGED)

__all__ = ["Solver", "SemError", "rect_quads", "baskets", "change_turn_quad", "quad2spectral", "extrude", "laminate_shell",
           "laminate_shell_delam", "boundary_nodes", "split_delam",
           "laminate_solid_smeared", "laminate_solid_layered", "lamina_q", "lamina_q3d", "SEM_SHELL5", "SEM_SOLID3",
           "SEM_FP64", "SEM_FP32_SPLIT", "SEM_Q_DISP", "SEM_Q_VEL", "SEM_Q_DISP_POST", "SEM_Q_ACC"]

PRECISIONS = {"fp64": SEM_FP64, "fp32_split": SEM_FP32_SPLIT}
QUANTITIES = {"disp": SEM_Q_DISP, "displacement": SEM_Q_DISP, "vel": SEM_Q_VEL, "velocity": SEM_Q_VEL,
              "disp_post": SEM_Q_DISP_POST, "acc": SEM_Q_ACC, "acceleration": SEM_Q_ACC}


def _check(rc, ctx=None):
    if rc != 0:
        lib = L.load()
        msg = lib.sem_last_error(ctx).decode() if ctx else ""
        raise SemError(rc, msg)


# ---------------------------------------------------------------------------- mesh
def rect_quads(lx, ly, nex, ney, x0=0.0, y0=0.0):
    """Structured bilinear quads over a rectangle: corner nodes row-major (x fastest),
    quads counter-clockwise -- the input ``quad2spectral`` expects (the reference gets
    it from gmsh's transfinite mesh, ``automesh_rectangle.m:47-57``)."""
    xs = x0 + lx * np.arange(nex + 1) / nex
    ys = y0 + ly * np.arange(ney + 1) / ney
    X, Y = np.meshgrid(xs, ys)
    pos = np.stack([X.reshape(-1), Y.reshape(-1), np.zeros(X.size)], axis=1)
    j, i = np.meshgrid(np.arange(ney), np.arange(nex), indexing="ij")
    n1 = (j * (nex + 1) + i).reshape(-1)
    quads = np.stack([n1, n1 + 1, n1 + nex + 2, n1 + nex + 1], axis=1)
    return quads.astype(np.int32), pos


def change_turn_quad(quads, pos):
    """Bilinear quads renumbered counter-clockwise where a corner's Jacobian determinant is not
    positive (``change_turn_quad.m``; the meshers call it before ``quad2spectral``,
    ``automesh_rectangle.m:57``).  Returns (quads, number renumbered)."""
    lib = L.load()
    q = np.array(L.i32(quads), copy=True)
    pos = L.f64(pos)
    if pos.shape[1] == 2:
        pos = L.f64(np.column_stack([pos, np.zeros(pos.shape[0])]))
    nf = C.c_int64(0)
    _check(lib.sem_mesh_change_turn_quad(q.shape[0], L.ptr(q, C.c_int32), pos.shape[0], L.ptr(pos, C.c_double),
                                         C.byref(nf)))
    return q, nf.value


def quad2spectral(quads, pos, order):
    """Spectral-element connectivity from bilinear quads (``quad2spectral_Fiborek.m``)."""
    lib = L.load()
    quads = L.i32(quads)
    pos = L.f64(pos)
    nq = quads.shape[0]
    n = order + 1
    ni = n - 2
    nmax = pos.shape[0] + nq * (4 * ni + ni * ni)
    nodes = np.empty((nq, n * n), np.int32)
    coords = np.empty((nmax, 3), np.float64)
    nn = C.c_int64(0)
    _check(lib.sem_mesh_quad2spectral(nq, L.ptr(quads, C.c_int32), pos.shape[0], L.ptr(pos, C.c_double),
                                      order, L.ptr(nodes, C.c_int32), L.ptr(coords, C.c_double),
                                      C.byref(nn)))
    return nodes, coords[: nn.value].copy()


def boundary_nodes(nodes, n_xy):
    """Sorted nodes on the mesh's boundary edges (``quad2spectral_Fiborek.m:153-156``)."""
    lib = L.load()
    nodes = L.i32(nodes)
    out = np.empty(nodes.shape[0] * 4 * n_xy, np.int32)
    nb = C.c_int64(0)
    _check(lib.sem_mesh_boundary_nodes(nodes.shape[0], n_xy, L.ptr(nodes, C.c_int32), L.ptr(out, C.c_int32),
                                       C.byref(nb)))
    return out[: nb.value].copy()


def split_delam(nodes, coords, den, n_xy, boundary=None):
    """Delamination by node split (``split_delam_nodes_flat_shell.m``): returns
    (nodes, coords, den_under, den_above) with the delaminated elements ``den`` appended
    again as the elements above the split interface."""
    lib = L.load()
    nodes = L.i32(nodes)
    coords = L.f64(coords)
    den = L.i32(np.atleast_1d(den))
    bnd = L.i32(boundary_nodes(nodes, n_xy) if boundary is None else boundary)
    n_el, nl = nodes.shape
    nodes_out = np.empty((n_el + den.size, nl), np.int32)
    coords_out = np.empty((coords.shape[0] + den.size * nl, 3), np.float64)
    under = np.empty(den.size, np.int32)
    above = np.empty(den.size, np.int32)
    nn = C.c_int64(0)
    _check(lib.sem_mesh_split_delam(n_el, n_xy, L.ptr(nodes, C.c_int32), coords.shape[0], L.ptr(coords, C.c_double),
                                    den.size, L.ptr(den, C.c_int32), bnd.size, L.ptr(bnd, C.c_int32),
                                    L.ptr(nodes_out, C.c_int32), L.ptr(coords_out, C.c_double), C.byref(nn),
                                    L.ptr(under, C.c_int32), L.ptr(above, C.c_int32)))
    return nodes_out, coords_out[: nn.value].copy(), under, above


def extrude(nodes2d, coords2d, n_xy, nz, h_layers):
    """3D mesh by extrusion (``extrude_mesh.m``): Nz = len(h_layers) element layers."""
    lib = L.load()
    nodes2d = L.i32(nodes2d)
    coords2d = L.f64(coords2d)
    h = L.f64(h_layers)
    Nz = h.size
    n2 = coords2d.shape[0]
    nodes3 = np.empty((Nz * nodes2d.shape[0], n_xy * n_xy * nz), np.int32)
    coords3 = np.empty((n2 * (Nz * (nz - 1) + 1), 3), np.float64)
    _check(lib.sem_mesh_extrude(nodes2d.shape[0], n_xy, L.ptr(nodes2d, C.c_int32), n2,
                                L.ptr(coords2d, C.c_double), nz, Nz, L.ptr(h, C.c_double),
                                L.ptr(nodes3, C.c_int32), L.ptr(coords3, C.c_double)))
    return nodes3, coords3


def baskets(nodes, n_nodes):
    """12-basket maps (``nodesMaps_Fiborek.m``), 1-based [rows x 12], 0 = empty."""
    lib = L.load()
    nodes = L.i32(nodes)
    n_el, n_loc = nodes.shape
    rows = -(-n_el * n_loc // 12)
    IG = np.zeros(12 * rows, np.int32)
    IL = np.zeros(12 * rows, np.int32)
    _check(lib.sem_mesh_baskets(n_el, n_loc, L.ptr(nodes, C.c_int32), int(n_nodes), L.ptr(IG, C.c_int32),
                                L.ptr(IL, C.c_int32)))
    return IG.reshape(12, rows).T, IL.reshape(12, rows).T


# ---------------------------------------------------------------------------- material
def lamina_q(E11, E22, G12, G13, G23, nu12):
    """Plane-stress reduced stiffness (``composite_plate.m:23-28``): q11 q12 q22 q44 q55 q66."""
    nu21 = E22 / E11 * nu12
    den = 1 - nu12 * nu21
    return np.array([E11 / den, nu12 * E22 / den, E22 / den, G23, G13, G12])


def lamina_q3d(E11, E22, E33, G12, G13, G23, nu12, nu13, nu23):
    """Orthotropic 3D stiffness (``composite3D_transf.m:31-46``): q11 q12 q13 q22 q23 q33 q44 q55 q66."""
    nu21 = E22 / E11 * nu12
    nu31 = E33 / E11 * nu13
    nu32 = E33 / E22 * nu23
    Delta = 1 - nu12 * nu21 - nu13 * nu31 - nu12 * nu23 * nu31 - nu13 * nu21 * nu32 - nu23 * nu32
    q11 = E11 * (1 - nu23 * nu32) / Delta
    q22 = E22 * (1 - nu31 * nu13) / Delta
    q33 = E33 * (1 - nu12 * nu21) / Delta
    q12 = (nu21 + nu31 * nu23) * E11 / Delta
    q13 = (nu31 + nu21 * nu32) * E11 / Delta
    q23 = (nu32 + nu12 * nu31) * E22 / Delta
    return np.array([q11, q12, q13, q22, q23, q33, G23, G13, G12])


def laminate_shell(lh, alpha_deg, rho, q):
    """23 shell coefficients of a laminate (``composite_plate_c_delam.m``); q is [lay x 6]
    or one row for all plies."""
    lib = L.load()
    lh = L.f64(lh)
    lay = lh.size
    al = L.f64(np.broadcast_to(alpha_deg, (lay,)))
    rh = L.f64(np.broadcast_to(rho, (lay,)))
    qq = L.f64(np.broadcast_to(np.asarray(q, float).reshape(-1, 6), (lay, 6)))
    out = np.empty(23)
    _check(lib.sem_laminate_shell(lay, L.ptr(lh, C.c_double), L.ptr(al, C.c_double), L.ptr(rh, C.c_double),
                                  L.ptr(qq, C.c_double), L.ptr(out, C.c_double)))
    return out


def laminate_shell_delam(lh, alpha_deg, rho, q, delamination_layer, side):
    """23 shell coefficients of the sub-laminate above (``side`` = +1: plies
    1..delamination_layer) or under (-1) a delamination (``composite_plate_c_delam.m:89-173``)."""
    lib = L.load()
    lh = L.f64(lh)
    lay = lh.size
    al = L.f64(np.broadcast_to(alpha_deg, (lay,)))
    rh = L.f64(np.broadcast_to(rho, (lay,)))
    qq = L.f64(np.broadcast_to(np.asarray(q, float).reshape(-1, 6), (lay, 6)))
    out = np.empty(23)
    _check(lib.sem_laminate_shell_delam(lay, L.ptr(lh, C.c_double), L.ptr(al, C.c_double), L.ptr(rh, C.c_double),
                                        L.ptr(qq, C.c_double), int(delamination_layer), int(side),
                                        L.ptr(out, C.c_double)))
    return out


def laminate_solid_smeared(lh, alpha_deg, rho, q3):
    """14 smeared solid coefficients (``composite3D_transf.m``, prop='smeared')."""
    lib = L.load()
    lh = L.f64(lh)
    lay = lh.size
    al = L.f64(np.broadcast_to(alpha_deg, (lay,)))
    rh = L.f64(np.broadcast_to(rho, (lay,)))
    qq = L.f64(np.broadcast_to(np.asarray(q3, float).reshape(-1, 9), (lay, 9)))
    out = np.empty(14)
    _check(lib.sem_laminate_solid_smeared(lay, L.ptr(lh, C.c_double), L.ptr(al, C.c_double),
                                          L.ptr(rh, C.c_double), L.ptr(qq, C.c_double), L.ptr(out, C.c_double)))
    return out


def laminate_solid_layered(alpha_deg, rho, q3):
    """[lay x 14] per-ply solid coefficient rows (``composite3D_transf.m:109-119``,
    prop='layered'): for meshes with one element layer per ply."""
    lib = L.load()
    al = L.f64(np.atleast_1d(np.asarray(alpha_deg, float)))
    lay = al.size
    rh = L.f64(np.broadcast_to(rho, (lay,)))
    qq = L.f64(np.broadcast_to(np.asarray(q3, float).reshape(-1, 9), (lay, 9)))
    out = np.empty((lay, 14))
    _check(lib.sem_laminate_solid_layered(lay, L.ptr(al, C.c_double), L.ptr(rh, C.c_double),
                                          L.ptr(qq, C.c_double), L.ptr(out, C.c_double)))
    return out


# ---------------------------------------------------------------------------- solver
class Solver:
    """One library context (one GPU, one stream)."""

    def __init__(self, device=0):
        self._lib = L.load()
        h = C.c_void_p()
        rc = self._lib.sem_create(C.byref(h), int(device))
        if rc != 0:
            raise SemError(rc, "sem_create failed (no HIP device?)")
        self._h = h
        self.kind = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _check(rc, self._h)

    # -- mesh
    def upload_mesh(self, kind, nodes, coords, n_xy, n_z=1):
        nodes = L.i32(nodes)
        coords = L.f64(coords)
        self._c(self._lib.sem_upload_mesh(self._h, kind, n_xy, n_z, nodes.shape[0], L.ptr(nodes, C.c_int32),
                                          coords.shape[0], L.ptr(coords, C.c_double)))
        self._mesh(kind, nodes.shape[0], coords.shape[0])

    def _mesh(self, kind, n_el, n_nodes):
        """Sizes of the mesh the library holds, recorded once an upload succeeded; the
        outputs and frames of an earlier mesh no longer apply."""
        self.kind = kind
        self.nf = 5 if kind == SEM_SHELL5 else 3
        self.n_el = n_el
        self.n_nodes = n_nodes
        self.dof = self.nf * n_nodes
        self.n_out = self.n_frames = 0  # the library dropped them with the old mesh

    def upload_batch(self, kind, cases, n_xy, n_z=1):
        """cases: list of (nodes [n_el x n_loc] with case-local ids, coords [n_nodes x 3]).
        Returns node offsets (case k's node v is union node off[k] + v)."""
        el_ptr = np.zeros(len(cases) + 1, np.int64)
        node_ptr = np.zeros(len(cases) + 1, np.int64)
        for k, (nd, co) in enumerate(cases):
            el_ptr[k + 1] = el_ptr[k] + np.asarray(nd).shape[0]
            node_ptr[k + 1] = node_ptr[k] + np.asarray(co).shape[0]
        nodes = L.i32(np.concatenate([np.asarray(nd) for nd, _ in cases]))
        coords = L.f64(np.concatenate([np.asarray(co) for _, co in cases]))
        self._c(self._lib.sem_upload_batch(self._h, kind, n_xy, n_z, len(cases), L.ptr(el_ptr, C.c_int64),
                                           L.ptr(nodes, C.c_int32), L.ptr(node_ptr, C.c_int64),
                                           L.ptr(coords, C.c_double)))
        self._mesh(kind, int(el_ptr[-1]), int(node_ptr[-1]))
        return node_ptr[:-1].copy(), el_ptr

    def baskets(self):
        rows = C.c_int64(0)
        self._c(self._lib.sem_basket_rows(self._h, C.byref(rows)))
        IG = np.zeros(12 * rows.value, np.int32)
        IL = np.zeros(12 * rows.value, np.int32)
        self._c(self._lib.sem_get_baskets(self._h, L.ptr(IG, C.c_int32), L.ptr(IL, C.c_int32)))
        return IG.reshape(12, rows.value).T, IL.reshape(12, rows.value).T  # [rows x 12], 1-based

    def is_affine(self):
        a = C.c_int(0)
        self._c(self._lib.sem_mesh_is_affine(self._h, C.byref(a)))
        return bool(a.value)

    def set_kernel(self, variant, patch_elems=0):
        """Element-kernel variant (0 point-wise, 1 line-wise) and the patch size used by the
        next upload_mesh (0 = the variant's default)."""
        self._c(self._lib.sem_set_kernel(self._h, int(variant), int(patch_elems)))

    def kernel(self):
        v = C.c_int(0); p = C.c_int(0)
        self._c(self._lib.sem_get_kernel(self._h, C.byref(v), C.byref(p)))
        return v.value, p.value

    # -- material
    def set_material(self, coef, elem_mat=None, stiff_scale=None, rho_scale=None):
        coef = L.f64(np.atleast_2d(coef))
        em = L.i32(np.zeros(self.n_el) if elem_mat is None else elem_mat)
        ss = None if stiff_scale is None else L.f64(stiff_scale)
        rs = None if rho_scale is None else L.f64(rho_scale)
        self._c(self._lib.sem_set_material(self._h, coef.shape[0], coef.shape[1], L.ptr(coef, C.c_double),
                                           L.ptr(em, C.c_int32), L.ptr(ss, C.c_double), L.ptr(rs, C.c_double)))

    def set_damping(self, dv_field):
        dv = L.f64(dv_field)
        self._c(self._lib.sem_set_damping(self._h, L.ptr(dv, C.c_double)))

    def mass(self):
        m = np.empty(self.dof)
        self._c(self._lib.sem_get_mass(self._h, L.ptr(m, C.c_double)))
        return m

    # -- excitation / outputs
    def set_excitation(self, dofs, amp, st, sig=None):
        """Point sources amp[k] at dofs[k] driven by st (1-based step nn reads st[nn-1]).
        With ``sig`` given, ``st`` is [n_sig x nft] and source k follows row sig[k]."""
        dofs = L.i64(np.atleast_1d(dofs))
        amp = L.f64(np.broadcast_to(amp, dofs.shape))
        if sig is None:
            st = L.f64(st)
            self._c(self._lib.sem_set_excitation(self._h, dofs.size, L.ptr(dofs, C.c_int64),
                                                 L.ptr(amp, C.c_double), st.size, L.ptr(st, C.c_double)))
            self.nft = st.size
            return
        st = L.f64(np.atleast_2d(st))
        sig = L.i32(np.broadcast_to(sig, dofs.shape))
        self._c(self._lib.sem_set_excitation_multi(self._h, dofs.size, L.ptr(dofs, C.c_int64),
                                                   L.ptr(amp, C.c_double), L.ptr(sig, C.c_int32), st.shape[0],
                                                   st.shape[1], L.ptr(st, C.c_double)))
        self.nft = st.shape[1]

    def set_outputs(self, dofs_or_functionals, quantity="disp"):
        """Point DOFs (list of ints) or functionals (list of (dofs, weights))."""
        ptr, dd, ww = [0], [], []
        for item in dofs_or_functionals:
            if np.isscalar(item):
                dd.append(int(item)); ww.append(1.0)
            else:
                ds, ws = item
                dd.extend(int(x) for x in ds); ww.extend(float(x) for x in ws)
            ptr.append(len(dd))
        n_out = len(ptr) - 1
        p = L.i64(ptr); d = L.i64(dd); w = L.f64(ww)
        q = QUANTITIES[quantity] if isinstance(quantity, str) else int(quantity)
        self._c(self._lib.sem_set_outputs(self._h, n_out, L.ptr(p, C.c_int64), L.ptr(d, C.c_int64),
                                          L.ptr(w, C.c_double), q))
        self.n_out = n_out  # only once the library took them: a refused call keeps the old outputs

    def set_frames(self, frame_nn, quantity="velocity"):
        f = L.i64(frame_nn)
        q = QUANTITIES[quantity] if isinstance(quantity, str) else int(quantity)
        self._c(self._lib.sem_set_frames(self._h, f.size, L.ptr(f, C.c_int64), q))
        self.n_frames = f.size  # only once the library took them

    def set_guard(self, limit, field=-1, post=False):
        self._c(self._lib.sem_set_guard(self._h, float(limit), int(field), int(bool(post))))

    # -- state
    def set_state(self, u=None, uold=None):
        u = None if u is None else L.f64(u)
        uo = None if uold is None else L.f64(uold)
        self._c(self._lib.sem_set_state(self._h, L.ptr(u, C.c_double), L.ptr(uo, C.c_double)))

    def get_state(self):
        u = np.empty(self.dof)
        uo = np.empty(self.dof)
        self._c(self._lib.sem_get_state(self._h, L.ptr(u, C.c_double), L.ptr(uo, C.c_double)))
        return u, uo

    # -- hot path
    def run(self, dt, nn_first, nn_last, precision="fp64", ld=None, want_frames=True, traces_out=None):
        """Runs steps nn_first..nn_last.  Returns dict(traces [ld x n_out], frames, t_frames,
        nn_done, seconds, status).  traces_out (optional): a C-contiguous float64 array
        [n_out x ld] the library writes the run's rows nn_first-1..nn_last-1 into (other rows
        are left as they are), e.g. one buffer shared by consecutive runs; traces is its
        transpose."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        ld = nn_last if ld is None else ld
        n_out = getattr(self, "n_out", 0)
        n_frames = getattr(self, "n_frames", 0)
        if traces_out is None:
            traces = np.zeros((n_out, ld))  # column-major [ld x n_out] == row-major [n_out x ld]
        else:
            traces = traces_out
            if (not isinstance(traces, np.ndarray) or traces.dtype != np.float64 or traces.shape != (n_out, ld)
                    or not traces.flags.c_contiguous or not traces.flags.writeable):
                raise ValueError(f"traces_out must be a writeable C-contiguous float64 array of shape ({n_out}, {ld})")
        frames = np.zeros((n_frames, self.dof)) if (n_frames and want_frames) else None
        tf = np.zeros(n_frames) if n_frames else None
        done = C.c_int64(0)
        secs = C.c_double(0)
        rc = self._lib.sem_run(self._h, float(dt), int(nn_first), int(nn_last), prec,
                               L.ptr(traces, C.c_double) if n_out else None, ld,
                               L.ptr(frames, C.c_double), L.ptr(tf, C.c_double), C.byref(done), C.byref(secs))
        status = "ok"
        if rc == SEM_E_DIVERGED:
            status = "diverged"
        else:
            self._c(rc)
        return dict(traces=traces.T, frames=frames, t_frames=tf, nn_done=done.value,
                    seconds=secs.value, status=status)

    def profile(self, dt, nn_first, nsteps, precision="fp64"):
        """Mean per-kernel durations [ms] (element, node, post-step) over nsteps steps."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        ms = np.zeros(3)
        self._c(self._lib.sem_profile(self._h, float(dt), int(nn_first), int(nsteps), prec,
                                      L.ptr(ms, C.c_double)))
        return ms

    def adjoint(self, dt, nn_first, nn_last, dJ, precision="fp64", seg_len=0):
        """Gradients (dJ/dstiff_scale, dJ/drho_scale) per element of a functional J of the
        traces of ``run(dt, nn_first, nn_last)`` from the current state, given dJ/dtraces
        as an array [nn_last x n_out] (row nn-1 per step).  Advances the state like run.
        Returns dict(g_stiff, g_rho, seconds)."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        n_out = getattr(self, "n_out", 0)
        ld = int(nn_last)
        dj = np.zeros((max(n_out, 1), ld))
        if n_out:
            a = np.asarray(dJ, dtype=np.float64)
            if a.ndim == 1:
                a = a[:, None]
            if a.shape[1] != n_out or a.shape[0] < nn_last:
                raise ValueError(f"dJ must be [>= {nn_last} x {n_out}], got {a.shape}")
            dj[:n_out] = a[:nn_last].T
        dj = L.f64(dj)
        gs = np.zeros(self.n_el)
        gr = np.zeros(self.n_el)
        sec = C.c_double(0.0)
        self._c(self._lib.sem_adjoint(self._h, float(dt), int(nn_first), int(nn_last), prec, L.ptr(dj, C.c_double),
                                      ld, L.ptr(gs, C.c_double), L.ptr(gr, C.c_double), int(seg_len),
                                      C.byref(sec)))
        return dict(g_stiff=gs, g_rho=gr, seconds=sec.value)

    def apply_K_noalloc(self, u_ptr, f_ptr, precision="fp64"):
        """``sem_apply_K_noalloc``: F_int = K u on device buffers without any allocation;
        ``u_ptr`` and ``f_ptr`` are device addresses (ints) of [dof] doubles on the context's
        device, in the caller's DOF order."""
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._c(self._lib.sem_apply_K_noalloc(self._h, C.c_void_p(int(u_ptr)), C.c_void_p(int(f_ptr)), prec))

    def apply_K(self, u, precision="fp64"):
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        u = L.f64(u)
        f = np.empty(self.dof)
        self._c(self._lib.sem_apply_K(self._h, L.ptr(u, C.c_double), L.ptr(f, C.c_double), prec))
        return f
