"""Wavefield on a uniform grid (``spec2meshgrid_flat_shell.m:1-583``) through the library's
``sem_grid_*`` entry points: the owner search and inverse mapping run once on the host,
the per-frame interpolation on the GPU."""
from __future__ import annotations

from pathlib import Path

import ctypes as C
import numpy as np

from . import _lib as L
from ._lib import SemError

__all__ = ["WavefieldGrid", "grid_plan", "spec2meshgrid_flat_shell", "variable_name", "MOTIONS"]

# flat_shell_variable_names.m: motion -> name suffix (prefix U/V/A by field variable)
MOTIONS = {1: "x", 2: "y", 3: "z", 4: "Fix", 5: "Fiy", 6: "x_top", 7: "y_top"}
_NAMES_8_9 = {"displacement": ("displacements_in_plane", "total_displacements"),
              "velocity": ("velocities_in_plane", "total_velocities"),
              "acceleration": ("accelerations_in_plane", "total_accelerations")}
_PREFIX = {"displacement": "U", "velocity": "V", "acceleration": "A"}


def variable_name(field_variable, motion):
    """``flat_shell_variable_names.m``."""
    if motion in (8, 9):
        return _NAMES_8_9[field_variable][motion - 8]
    return _PREFIX[field_variable] + MOTIONS[motion]


def grid_plan(nodes, coords, n_xy, Nx, Ny, n_search=0):
    """The host set-up alone (no GPU): grid coordinates, owner element and local
    coordinates of every point (``sem_grid_plan``)."""
    lib = L.load()
    nodes = L.i32(nodes)
    coords = L.f64(coords)
    P = Nx * Ny
    xi, yi = np.empty(Nx), np.empty(Ny)
    owner = np.empty(P, np.int32)
    s, t = np.empty(P), np.empty(P)
    rc = lib.sem_grid_plan(nodes.shape[0], n_xy, L.ptr(nodes, C.c_int32), coords.shape[0], L.ptr(coords, C.c_double),
                           int(n_search), Nx, Ny, L.ptr(xi, C.c_double), L.ptr(yi, C.c_double),
                           L.ptr(owner, C.c_int32), L.ptr(s, C.c_double), L.ptr(t, C.c_double))
    if rc != 0:
        raise SemError(rc, "grid set-up failed (a grid point outside every searched element?)")
    return dict(xi=xi, yi=yi, owner=owner, s=s, t=t)


class WavefieldGrid:
    """Nx x Ny grid over a flat-shell mesh; ``n_search`` leading elements are searched
    for owners (0 = all; pass the element count before the delaminated copies)."""

    def __init__(self, nodes, coords, n_xy, Nx, Ny, n_search=0, device=0):
        self._lib = L.load()
        nodes = L.i32(nodes)
        coords = L.f64(coords)
        if coords.shape[1] == 2:
            coords = L.f64(np.concatenate([coords, np.zeros((coords.shape[0], 1))], axis=1))
        self.n_nodes = coords.shape[0]
        self.Nx, self.Ny = int(Nx), int(Ny)
        h = C.c_void_p()
        rc = self._lib.sem_grid_create(C.byref(h), device, nodes.shape[0], n_xy, L.ptr(nodes, C.c_int32),
                                       self.n_nodes, L.ptr(coords, C.c_double), int(n_search), self.Nx, self.Ny)
        if rc != 0:
            raise SemError(rc, self._lib.sem_grid_last_error(None).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sem_grid_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def points(self):
        P = self.Nx * self.Ny
        xi, yi = np.empty(self.Nx), np.empty(self.Ny)
        owner = np.empty(P, np.int32)
        s, t = np.empty(P), np.empty(P)
        self._check(self._lib.sem_grid_points(self._h, L.ptr(xi, C.c_double), L.ptr(yi, C.c_double),
                                              L.ptr(owner, C.c_int32), L.ptr(s, C.c_double), L.ptr(t, C.c_double)))
        return dict(xi=xi, yi=yi, owner=owner, s=s, t=t)

    def eval(self, frames, motion, surface="top", h=0.0, twin_under=None, twin_above=None):
        """[n_frames, Ny, Nx] of quantity ``motion`` (1-9) from frames [n_frames, dof]."""
        fr = L.f64(np.atleast_2d(frames))
        tu = L.i32(np.zeros(0) if twin_under is None else np.asarray(twin_under).reshape(-1))
        ta = L.i32(np.zeros(0) if twin_above is None else np.asarray(twin_above).reshape(-1))
        if tu.size != ta.size:
            raise ValueError("twin lists differ in length")
        out = np.empty((fr.shape[0], self.Ny, self.Nx))
        self._check(self._lib.sem_grid_eval(self._h, fr.shape[0], fr.shape[1], L.ptr(fr, C.c_double), int(motion),
                                            1 if surface == "top" else -1, float(h), tu.size,
                                            L.ptr(tu, C.c_int32), L.ptr(ta, C.c_int32), L.ptr(out, C.c_double)))
        return out

    def _check(self, rc):
        if rc != 0:
            raise SemError(rc, self._lib.sem_grid_last_error(self._h).decode())


def spec2meshgrid_flat_shell(test_case, mesh, Nx, Ny, field_variable, motion, shell_surface, h, frames=None,
                             input_name=None, n_frames=None, output_name=None, device=0):
    """Mirror of ``spec2meshgrid_flat_shell.m``: ``mesh`` is a dict as ``matio.load_mesh``
    returns (or a mesh file path); frames come as an array [n_frames, dof] or are read
    from the frame files under ``input_name``.  Returns Data [Ny, Nx, n_frames] (the
    reference's layout) and, with ``output_name``, writes
    ``flat_shell_<variable>_<test_case>_<Nx>x<Ny><surface>.mat`` {Data}."""
    from . import matio
    if not isinstance(mesh, dict):
        mesh = matio.load_mesh(mesh)
    nodes, coords = mesh["nodes"], mesh["coords"]
    n_xy = int(round(np.sqrt(nodes.shape[1])))
    ups, dns = mesh.get("den_above", []), mesh.get("den_under", [])
    n_copies = sum(len(a) for a in (ups if shell_surface == "top" else dns))
    tu = ta = None
    if shell_surface == "top" and ups:
        tu = np.concatenate([nodes[d].reshape(-1) for d in dns])
        ta = np.concatenate([nodes[a].reshape(-1) for a in ups])
    if frames is None:
        frames = _read_frames(input_name, field_variable, motion, n_frames, coords.shape[0])
    with WavefieldGrid(nodes, coords, n_xy, Nx, Ny, n_search=nodes.shape[0] - n_copies, device=device) as g:
        data = g.eval(frames, motion, shell_surface, h, tu, ta)
    Data = np.transpose(data, (1, 2, 0))  # [Ny, Nx, n_frames]
    if output_name is not None:
        import scipy.io as sio
        out = Path(output_name)
        out.mkdir(parents=True, exist_ok=True)
        name = f"flat_shell_{variable_name(field_variable, motion)}_{test_case}_{Nx}x{Ny}{shell_surface}.mat"
        sio.savemat(str(out / name), {"Data": Data})
    return Data


def _read_frames(input_name, field_variable, motion, n_frames, n_nodes):
    """Frames [n_frames, 5 n_nodes] from the reference's frame files (the fields the
    quantity needs; the others stay zero)."""
    import scipy.io as sio
    pre = _PREFIX[field_variable]
    need = {1: ("x",), 2: ("y",), 3: ("z",), 4: ("Fix",), 5: ("Fiy",), 6: ("x", "Fix"), 7: ("y", "Fiy"),
            8: ("x", "Fix", "y", "Fiy"), 9: ("x", "Fix", "y", "Fiy", "z")}[motion]
    col = {"x": 0, "Fix": 1, "y": 2, "Fiy": 3, "z": 4}
    out = np.zeros((n_frames, n_nodes, 5))
    for c in range(n_frames):
        for nm in need:
            v = pre + nm
            out[c, :, col[nm]] = sio.loadmat(str(Path(input_name) / f"{v}_frame{c + 1}.mat"))[v].reshape(-1)
    return out.reshape(n_frames, -1)
