"""The reference's on-disk formats, so that MATLAB callers and post-processing scripts
read and write the same files.

* Mesh file, as the meshers save it and the engines ``load(meshfile)``
  (``mesher/automesh_simple.m:105``): ``nodes`` [fen x n^2] and ``coords`` [N x 3] with
  1-based node ids, ``I_G``/``I_L`` (12-basket maps, 1-based), ``pztEl``, ``delamEl``,
  ``den_under``/``den_above`` (cells of 1-based element lists), ``mesh_min``/``mesh_max``,
  ``shape_order``.
* Engine outputs (``main_flat_shell_force_excit_c.m:256-258, 742-877``): ``displ<k>.mat``
  {displ [nft x n_out]}, ``voltage<k>.mat`` {voltage}, ``time<k>.mat`` {minTime,
  averageTime, TotalNofNodes, t}, and per frame c the five field files
  ``{U|V|A}{x,Fix,y,Fiy,z}_frame<c>.mat``, each holding that field of all nodes.

Files are MAT v5 (``scipy.io``), which MATLAB reads and writes; loading parses data only.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy.io as sio

__all__ = ["load_mesh", "save_mesh", "save_outputs", "save_frames", "FRAME_PREFIX"]

# field_variable -> file prefixes (main_flat_shell_force_excit_c.m:742-842)
FRAME_PREFIX = {"displacement": ("U",), "velocity": ("V",), "acceleration": ("A",), "all": ("U", "V", "A")}
_SHELL_FIELDS = ("x", "Fix", "y", "Fiy", "z")  # u, phi_x, v, phi_y, w


def _cells(v):
    """A MATLAB cell of element lists (or one list) -> list of 0-based int arrays."""
    if v is None:
        return []
    a = np.asarray(v, dtype=object)
    if a.dtype != object:
        return [np.asarray(v, dtype=np.int64).reshape(-1) - 1] if np.size(v) else []
    return [np.asarray(x, dtype=np.int64).reshape(-1) - 1 for x in a.reshape(-1) if np.size(x)]


def load_mesh(path) -> dict:
    """Read a mesh file; node and element ids come back 0-based (``nodes`` int32)."""
    m = sio.loadmat(str(path), squeeze_me=False, mat_dtype=False)
    out = dict(nodes=np.asarray(m["nodes"], dtype=np.int64).astype(np.int32) - 1,
               coords=np.asarray(m["coords"], dtype=np.float64))
    if out["coords"].shape[1] == 2:
        out["coords"] = np.concatenate([out["coords"], np.zeros((out["coords"].shape[0], 1))], axis=1)
    if "shape_order" in m:
        out["shape_order"] = int(np.asarray(m["shape_order"]).reshape(-1)[0])
    for k in ("I_G", "I_L"):
        if k in m:
            out[k] = np.asarray(m[k], dtype=np.int64)
    for k in ("mesh_min", "mesh_max"):
        if k in m:
            out[k] = np.asarray(m[k], dtype=np.float64).reshape(-1)
    for k in ("pztEl", "delamEl", "den_under", "den_above"):
        if k in m:
            out[k] = _cells(m[k])
    return out


def _cell(lists):
    c = np.empty((1, len(lists)), dtype=object)
    for i, x in enumerate(lists):
        c[0, i] = np.asarray(x, dtype=np.float64).reshape(1, -1) + 1
    return c


def save_mesh(path, nodes, coords, shape_order, I_G=None, I_L=None, pztEl=(), delamEl=(), den_under=(),
              den_above=(), mesh_min=None, mesh_max=None):
    """Write a mesh file (0-based ids in, 1-based doubles on disk, as the meshers save)."""
    coords = np.asarray(coords, dtype=np.float64)
    d = dict(nodes=np.asarray(nodes, dtype=np.float64) + 1, coords=coords,
             shape_order=float(shape_order),
             mesh_min=np.asarray(coords.min(axis=0) if mesh_min is None else mesh_min, float).reshape(1, -1),
             mesh_max=np.asarray(coords.max(axis=0) if mesh_max is None else mesh_max, float).reshape(1, -1),
             pztEl=_cell(pztEl), delamEl=np.asarray(np.concatenate([np.asarray(x).reshape(-1) for x in delamEl])
                                                     if len(delamEl) else [], float).reshape(1, -1) + 1,
             den_under=_cell(den_under), den_above=_cell(den_above))
    if I_G is not None:
        d["I_G"] = np.asarray(I_G, dtype=np.float64)
    if I_L is not None:
        d["I_L"] = np.asarray(I_L, dtype=np.float64)
    sio.savemat(str(path), d, do_compression=False, oned_as="column")


def save_frames(output_name, frames, field_variable, nf=5, start=1):
    """Frame files of one engine run: ``frames`` [n_frames x dof] (caller's DOF order) for
    one quantity, written under the prefix of ``field_variable`` ('displacement',
    'velocity' or 'acceleration'), frame numbers from ``start``."""
    out = Path(output_name)
    out.mkdir(parents=True, exist_ok=True)
    (prefix,) = FRAME_PREFIX[field_variable]
    names = _SHELL_FIELDS if nf == 5 else ("x", "y", "z")
    for c, fr in enumerate(np.asarray(frames, dtype=np.float64)):
        for f, nm in enumerate(names):
            var = prefix + nm
            sio.savemat(str(out / f"{var}_frame{c + start}.mat"), {var: fr[f::nf].reshape(-1, 1)})


def save_outputs(output_name, test_case, displ=None, voltage=None, time=None):
    """``displ<k>.mat``, ``voltage<k>.mat`` and ``time<k>.mat`` of one engine run."""
    out = Path(output_name)
    out.mkdir(parents=True, exist_ok=True)
    if displ is not None:
        sio.savemat(str(out / f"displ{test_case}.mat"), {"displ": np.asarray(displ, dtype=np.float64)})
    if voltage is not None:
        sio.savemat(str(out / f"voltage{test_case}.mat"), {"voltage": np.asarray(voltage, dtype=np.float64)})
    if time is not None:
        sio.savemat(str(out / f"time{test_case}.mat"),
                    {"minTime": float(time["minTime"]), "averageTime": float(time["averageTime"]),
                     "TotalNofNodes": float(time["TotalNofNodes"]),
                     "t": np.asarray(time["t"], dtype=np.float64).reshape(1, -1)})
