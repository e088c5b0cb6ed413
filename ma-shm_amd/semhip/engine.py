"""Host-side mirror of the reference's solver entry points.

``main_flat_shell_force_excit_c`` follows ``main_flat_shell_force_excit_c.m:1-888``:
same inputs (the input-script workspace as a dict, the mesh arrays) and outputs
(``displ`` [nft x n_out], ``t_frames``, frames, ``time`` = minTime/averageTime/
TotalNofNodes/t), with the time loop executed by the HIP library.  ``solid3d_engine``
does the same for ``main_spec3Dcuda_global_cpu_gpu3_wave_sens_opt_single_PCT_mat.m``.
Case builders for the structured-plate workloads live here too.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import (Solver, SEM_SHELL5, SEM_SOLID3, rect_quads, quad2spectral, extrude, laminate_shell,
               laminate_shell_delam, split_delam,
               laminate_solid_smeared, laminate_solid_layered, lamina_q, lamina_q3d)


# ---------------------------------------------------------------------------- signal
def hanning_signal(dt, nft, f_1, f_2, t_1):
    """``Hanning_signal.m:17-23``."""
    t_2 = t_1 + 1.0 / f_1
    t = np.arange(nft, dtype=np.float64) * dt
    st = np.zeros(nft)
    m = (t >= t_1) & (t <= t_2)
    tm = t[m]
    st[m] = 0.5 * (1.0 - np.cos(2 * np.pi * f_1 * (tm - t_2))) * np.sin(2 * np.pi * f_2 * (tm - t_1))
    return t, st


def nearest_node(coords, xy):
    """``force_excit_c.m:63-72``: nearest node by Euclidean distance, first index wins."""
    d = np.sqrt((coords[:, 0] - xy[0]) ** 2 + (coords[:, 1] - xy[1]) ** 2)
    return int(np.argmin(d))


# ---------------------------------------------------------------------------- flat shell
@dataclass
class FlatShellCase:
    """A structured flat-shell plate with a transverse point force (config-1 style)."""
    lx: float = 0.5
    ly: float = 0.5
    nex: int = 40
    ney: int = 40
    order: int = 5
    lh: tuple = (0.25e-3,) * 8
    angles: tuple = (0, 90, 0, 90, 90, 0, 90, 0)
    E11: float = 130e9
    E22: float = 10e9
    G12: float = 5e9
    G13: float = 5e9
    G23: float = 3.5e9
    nu12: float = 0.3
    rho: float = 1570.0
    force_xy: tuple = (0.25, 0.25)
    V0: float = 1.0
    f_2: float = 50e3
    cycles: float = 5.0
    t_1: float = 0.0
    dt: float = 1e-7
    nft: int = 5000
    sensors: tuple = ((0.40, 0.25), (0.25, 0.40), (0.10, 0.25), (0.25, 0.10))
    out_quantity: str = "vel"
    n_frames: int = 10
    field_variable: str = "velocity"
    etad_xy: float = 0.0
    etad_z: float = 0.0
    stiff_scale: np.ndarray | None = None
    rho_scale: np.ndarray | None = None
    # transducers: "point" = transverse point force at force_xy and w outputs at the
    # sensor points; "pzt" = square PZT patches (side pzt_size, thickness pzt_t) centred at
    # pzt_xy, laminated into the plate; patch actuator_no is driven with V0 volts, every
    # patch reads a voltage (main_flat_shell_multi_pzt_c.m)
    transducers: str = "point"
    pzt_xy: tuple = ()
    actuator_no: int = 0
    pzt_size: float = 0.01
    pzt_t: float = 0.5e-3
    pzt_nz: int = 3
    # delamination (automesh_simple.m:86-98): the elements whose centroid lies in
    # delam_rect = (x0, y0, x1, y1) are split after ply delamination_layer
    # (split_delam_nodes_flat_shell.m); the copies above the interface take plies
    # 1..delamination_layer, the originals under it the others (composite_plate_c_delam.m:89-173)
    delam_rect: tuple = ()
    delamination_layer: int = 4
    # a mesh file as the meshers write it (matio.load_mesh: nodes, coords, den_under,
    # den_above, ...) in place of the structured nex x ney grid; its delaminations (if any)
    # take delamination_layer (one value, or one per delamination)
    mesh_file: str | None = None

    @property
    def dof(self):
        """DOF of the pristine mesh (a delamination adds its split nodes)."""
        n = self.order
        return 5 * (self.nex * n + 1) * (self.ney * n + 1)


def shell_plate_40x40(**kw):
    """0.5x0.5 m, 2 mm, [0/90/0/90]s CFRP, 40x40 order-5 shells, 50 kHz 5-cycle burst,
    dt = 1e-7 s, 5000 steps, w-velocity at 4 sensors on a 0.15 m circle, 10 frames."""
    return FlatShellCase(**kw)


def prepare_flat_shell(case: FlatShellCase):
    q = lamina_q(case.E11, case.E22, case.G12, case.G13, case.G23, case.nu12)
    coef = laminate_shell(np.asarray(case.lh), np.asarray(case.angles, float), case.rho, q)
    delam = None
    if case.mesh_file:
        from . import matio
        m = matio.load_mesh(case.mesh_file)
        nodes, coords = m["nodes"], m["coords"]
        if m.get("shape_order", case.order) != case.order or nodes.shape[1] != (case.order + 1) ** 2:
            raise ValueError("mesh file order differs from the case's order")
        ups, dns = m.get("den_above", []), m.get("den_under", [])
        if ups:
            lays = np.broadcast_to(np.asarray(case.delamination_layer), (len(ups),))
            lh, al = np.asarray(case.lh), np.asarray(case.angles, float)
            rows = [coef]
            em = np.zeros(nodes.shape[0], np.int32)
            for j, (a, b) in enumerate(zip(ups, dns)):  # composite_plate_c_delam.m:89-173
                rows += [laminate_shell_delam(lh, al, case.rho, q, int(lays[j]), +1),
                         laminate_shell_delam(lh, al, case.rho, q, int(lays[j]), -1)]
                em[a] = 2 * j + 1
                em[b] = 2 * j + 2
            coef = np.stack(rows)
            delam = dict(den_under=np.concatenate(dns), den_above=np.concatenate(ups), elem_mat=em)
    else:
        quads, pos = rect_quads(case.lx, case.ly, case.nex, case.ney)
        nodes, coords = quad2spectral(quads, pos, case.order)
    if delam and case.transducers != "point":
        raise ValueError("a delamination is modelled with the point-force transducers only")
    if case.delam_rect:
        if case.mesh_file:
            raise ValueError("delam_rect applies to the structured grid, not to a mesh file")
        if case.transducers != "point":
            raise ValueError("a delamination is modelled with the point-force transducers only")
        x0, y0, x1, y1 = case.delam_rect
        cen = coords[nodes].mean(axis=1)
        den = np.nonzero((cen[:, 0] > x0) & (cen[:, 0] < x1) & (cen[:, 1] > y0) & (cen[:, 1] < y1))[0]
        if den.size == 0:
            raise ValueError("delam_rect holds no element centroid")
        nodes, coords, den_under, den_above = split_delam(nodes, coords, den, case.order + 1)
        lh, al = np.asarray(case.lh), np.asarray(case.angles, float)
        coef = np.stack([coef, laminate_shell_delam(lh, al, case.rho, q, case.delamination_layer, +1),
                         laminate_shell_delam(lh, al, case.rho, q, case.delamination_layer, -1)])
        em = np.zeros(nodes.shape[0], np.int32)
        em[den_above] = 1
        em[den_under] = 2
        delam = dict(den_under=den_under, den_above=den_above, elem_mat=em)
    J = nearest_node(coords, case.force_xy)
    sens = [5 * nearest_node(coords, s) + 4 for s in case.sensors]
    f_1 = case.f_2 / case.cycles
    _, st = hanning_signal(case.dt, case.nft, f_1, case.f_2, case.t_1)
    frm_int = case.nft // case.n_frames if case.n_frames else 0
    frame_no = [frm_int * (c + 1) for c in range(case.n_frames)]
    h = float(np.sum(case.lh))
    out = dict(nodes=nodes, coords=coords, coef=coef, force_dof=5 * J + 4, sensor_dofs=sens, st=st,
               frame_no=frame_no, h=h, elem_mat=None, ex_dofs=np.array([5 * J + 4]), ex_amp=np.array([case.V0]))
    if delam:
        out.update(elem_mat=delam["elem_mat"], den_under=delam["den_under"], den_above=delam["den_above"])
    if case.transducers == "pzt":
        from . import pzt
        S, d, epsT, rho_p = pzt.pzt5a_strain_charge()
        Q, e, g, (e31, e32, g33) = pzt.pzt_const(S, d, epsT, "shell", 0.0)
        cen = coords[nodes].mean(axis=1)
        elems = [np.nonzero((np.abs(cen[:, 0] - x) < case.pzt_size / 2) & (np.abs(cen[:, 1] - y) < case.pzt_size / 2))[0]
                 for x, y in case.pzt_xy]
        em = np.zeros(nodes.shape[0], np.int32)
        for el in elems:
            em[el] = 1
        row = np.asarray(coef).reshape(-1)
        coef = np.stack([row, pzt.laminate_shell_add_pzt(row, h, case.pzt_t, rho_p, Q)])
        dofs, amp = pzt.shell_actuator(nodes[elems[case.actuator_no]], coords, case.order + 1, case.pzt_nz,
                                       case.pzt_t, h, e, g, case.V0)
        out.update(coef=coef, elem_mat=em, ex_dofs=dofs, ex_amp=amp, pzt_elems=elems,
                   pzt_consts=(e31, e32, g33))
    return out


# Element kernel (variant, patch elements) of the engine functions unless told otherwise:
# the line-wise kernel with 8-element patches, the autotuner's pick or within about 1% of
# it on every single-case workload measured (shells and solids, both forms).  None selects
# the library default (variant 0).
DEFAULT_KERNEL = (1, 8)


def flat_shell_solver(case: FlatShellCase, prep=None, device=0, kernel=DEFAULT_KERNEL):
    """Configured Solver for a flat-shell case; kernel = (variant, patch elements) or None
    for the library default."""
    prep = prepare_flat_shell(case) if prep is None else prep
    s = Solver(device)
    if kernel is not None:
        s.set_kernel(*kernel)
    s.upload_mesh(SEM_SHELL5, prep["nodes"], prep["coords"], case.order + 1)
    s.set_material(prep["coef"], elem_mat=prep["elem_mat"], stiff_scale=case.stiff_scale, rho_scale=case.rho_scale)
    s.set_damping([case.etad_xy, 0.0, case.etad_xy, 0.0, case.etad_z])
    s.set_excitation(prep["ex_dofs"], prep["ex_amp"], prep["st"])
    if case.transducers == "pzt":
        from . import pzt
        e31, e32, g33 = prep["pzt_consts"]
        prep["outputs"] = [pzt.shell_sensor(s, el, case.pzt_t, prep["h"], e31, e32, g33) for el in prep["pzt_elems"]]
        s.set_outputs(prep["outputs"], "disp")
    else:
        s.set_outputs(prep["sensor_dofs"], case.out_quantity)
    if prep["frame_no"]:
        s.set_frames(prep["frame_no"], case.field_variable)
    s.set_guard(10 * prep["h"], field=4, post=False)
    return s, prep


def main_flat_shell_force_excit_c(case: FlatShellCase, precision="fp64", device=0, output_name=None,
                                  test_case=1):
    """Mirror of ``main_flat_shell_force_excit_c.m``: returns displ, t_frames, frames and
    the timing record (minTime, averageTime, TotalNofNodes, t) for mode='hip'.  With
    ``output_name`` it also writes the reference's files there (``displ<k>.mat`` or, for
    PZT sensors, ``voltage<k>.mat``; ``time<k>.mat``; the frame files; ``:742-877``)."""
    s, prep = flat_shell_solver(case, device=device)
    t, _ = hanning_signal(case.dt, case.nft, case.f_2 / case.cycles, case.f_2, case.t_1)
    r = s.run(case.dt, 1, case.nft, precision)
    s.close()
    avg = r["seconds"] / max(case.nft - 1, 1)
    out = dict(displ=r["traces"], t_frames=r["t_frames"], frames=r["frames"], status=r["status"],
               nn_done=r["nn_done"], time=dict(minTime=avg, averageTime=avg,
                                              TotalNofNodes=int(prep["coords"].shape[0]), t=t),
               seconds=r["seconds"])
    if output_name is not None:
        from . import matio
        pzt = case.transducers == "pzt"
        matio.save_outputs(output_name, test_case, displ=None if pzt else out["displ"],
                           voltage=out["displ"] if pzt else None, time=out["time"])
        if len(out["frames"]):
            matio.save_frames(output_name, out["frames"], case.field_variable, nf=5)
    return out


# ---------------------------------------------------------------------------- 3D solid
@dataclass
class SolidCase:
    """Structured 3D laminated plate of hexahedral GLL elements (config-2/5 style)."""
    lx: float = 0.5
    ly: float = 0.5
    nex: int = 100
    ney: int = 100
    nlay_el: int = 2
    h: float = 2e-3
    order: int = 4
    order_z: int = 4
    lh: tuple = (0.25e-3,) * 8
    angles: tuple = (0, 90, 0, 90, 90, 0, 90, 0)
    E11: float = 130e9
    E22: float = 10e9
    G12: float = 5e9
    G13: float = 5e9
    G23: float = 3.5e9
    nu12: float = 0.3
    rho: float = 1570.0
    f_2: float = 100e3
    cycles: float = 5.0
    dt: float = 2e-8
    nft: int = 20000
    force_xy: tuple = (0.25, 0.25)
    V0: float = 1.0
    sensors: tuple = ((0.40, 0.25), (0.25, 0.40), (0.10, 0.25), (0.25, 0.10))
    out_quantity: str = "disp_post"
    n_frames: int = 0
    etad_xy: float = 0.0
    etad_z: float = 0.0
    stiff_scale: np.ndarray | None = None
    # transducers: "point" = transverse point force at force_xy (amplitude V0 N) and
    # w-displacement at the sensor points; "pzt" = square PZT-5A transducers of side
    # pzt_size and thickness pzt_t bonded on the top surface: the actuator at force_xy
    # driven with V0 volts, voltage sensors centred at the sensor points
    transducers: str = "point"
    pzt_size: float = 0.01
    pzt_t: float = 0.2e-3
    # with the PZT actuator: "pzt" = voltage sensors (PZT patches at the sensor points),
    # "point" = w-displacement point sensors on the top surface
    sensor_kind: str = "pzt"
    # where the transducers live: "surface" = bonded on the top surface, their forces and
    # voltage weights carried down to the plate nodes (pzt.surface_actuator); "meshed" = the
    # reference's layout (main_spec3Dcuda_global_..._mat.m:198-388, 499-537): the transducer
    # hexahedra are elements of the mesh, stacked on the plate's top faces with their own
    # nodes and the PZT material row, and the time step must suit their thin elements
    pzt_layout: str = "surface"
    # laminate stiffness (composite3D_transf.m:78-119): "smeared" = one layup-averaged row
    # for all plate elements; "layered" = one row per ply, for meshes with one element layer
    # per ply (nlay_el == len(lh)): element layer k from the bottom (the extrusion's element
    # order) takes ply k
    prop: str = "smeared"

    @property
    def dof(self):
        nxy = self.nex * self.order + 1
        return 3 * nxy * (self.ney * self.order + 1) * (self.nlay_el * self.order_z + 1)


PZT_CIRCLE = tuple((0.25 + 0.15 * np.cos(a), 0.25 + 0.15 * np.sin(a)) for a in np.arange(8) * np.pi / 4)


def solid_plate_100x100_pzt(**kw):
    """0.5x0.5x0.002 m [0/90/0/90]s plate, 100x100x2 hexahedra of order 4 (5x5x5 GLL),
    dt = 2e-8 s, 20000 steps; a 10 mm x 0.2 mm PZT-5A actuator at the centre driven by a
    100 kHz 5-cycle burst of 10 V, 8 PZT voltage sensors on a 0.15 m circle."""
    base = dict(transducers="pzt", V0=10.0, sensors=PZT_CIRCLE)
    base.update(kw)
    return SolidCase(**base)


SURFACE_GRID_64 = tuple((0.0625 + 0.125 * i, 0.0625 + 0.125 * j) for j in range(8) for i in range(8))


def solid_plate_250x250_pzt(**kw):
    """1.0x1.0x0.004 m [0/45/-45/90]2s plate, 250x250x2 hexahedra of order 5; a 10 mm PZT-5A
    actuator at the centre driven by a 75 kHz 5-cycle burst of 10 V; w displacement at 64
    surface points on an 8 x 8 grid; 20 snapshots over 30000 steps of 2e-8 s."""
    base = dict(lx=1.0, ly=1.0, nex=250, ney=250, nlay_el=2, h=4e-3, order=5, order_z=5,
                lh=(0.25e-3,) * 16, angles=(0, 45, -45, 90, 0, 45, -45, 90, 90, -45, 45, 0, 90, -45, 45, 0),
                f_2=75e3, nft=30000, force_xy=(0.5, 0.5), transducers="pzt", V0=10.0, sensor_kind="point",
                sensors=SURFACE_GRID_64, n_frames=20)
    base.update(kw)
    return SolidCase(**base)


def prepare_solid(case: SolidCase):
    quads, pos = rect_quads(case.lx, case.ly, case.nex, case.ney)
    n2, c2 = quad2spectral(quads, pos, case.order)
    hl = [case.h / case.nlay_el] * case.nlay_el
    nodes, coords = extrude(n2, c2, case.order + 1, case.order_z + 1, hl)
    if case.nlay_el > 2:
        # extrude_mesh.m:84 sets hlay = dz instead of adding it up, so every layer after the
        # second lands on the second and their shared faces collapse; with more than two
        # element layers (one per ply for prop="layered") layer k sits on the sum of the
        # layers below it.  Node ids are those of the extrusion: level = id // (2D nodes).
        from .pzt import gll_nodes
        ksi, _ = gll_nodes(case.order_z + 1)
        nzl = case.order_z
        base = np.concatenate([[0.0], np.cumsum(hl)])
        lev = np.arange(case.nlay_el * nzl + 1)
        lay = np.minimum(lev // nzl, case.nlay_el - 1)
        zlev = base[lay] + np.asarray(hl)[lay] * (ksi[lev - lay * nzl] + 1) / 2
        coords = coords.copy()
        coords[:, 2] = zlev[np.arange(coords.shape[0]) // c2.shape[0]]
    E33 = case.E22
    nu23 = case.E22 / (2 * case.G23) - 1.0  # transversely isotropic ply: G23 = E22 / (2 (1 + nu23))
    q3 = lamina_q3d(case.E11, case.E22, E33, case.G12, case.G13, case.G23, case.nu12, case.nu12, nu23)
    elem_mat = None
    if case.prop == "layered":
        if len(case.lh) != case.nlay_el:
            raise ValueError(f"layered stiffness needs one element layer per ply ({len(case.lh)} plies, "
                             f"{case.nlay_el} element layers)")
        coef = laminate_solid_layered(np.asarray(case.angles, float), case.rho, q3)
        elem_mat = (np.arange(nodes.shape[0]) // (case.nex * case.ney)).astype(np.int32)
    elif case.prop == "smeared":
        coef = laminate_solid_smeared(np.asarray(case.lh), np.asarray(case.angles, float), case.rho, q3)[None, :]
    else:
        raise ValueError(f"unknown laminate stiffness {case.prop!r}")
    top = coords[:, 2].max()
    topmask = np.isclose(coords[:, 2], top)
    idx_top = np.nonzero(topmask)[0]
    _, st = hanning_signal(case.dt, case.nft, case.f_2 / case.cycles, case.f_2, 0.0)
    frm_int = case.nft // case.n_frames if case.n_frames else 0
    frame_no = [frm_int * (c + 1) for c in range(case.n_frames)]
    out = dict(nodes=nodes, coords=coords, coef=coef, st=st, frame_no=frame_no)
    if elem_mat is not None:
        out["elem_mat"] = elem_mat
    if case.transducers == "pzt" and case.pzt_layout == "meshed":
        from . import pzt
        Q, e, g, rho_p = pzt.pzt5a()
        n, nz = case.order + 1, case.order_z + 1
        hx = case.lx / case.nex

        def footprint(xy):
            c = (round(xy[0] / hx) * hx, round(xy[1] / hx) * hx)  # centred on an element corner
            return pzt.top_elements(nodes, coords, c, case.pzt_size / 2, n, nz)

        # transducer hexahedra on the plate's top faces (their bottom level shares the plate's
        # top-face nodes), one transducer after the other; the PZT material row for them
        # (main_spec3Dcuda_global_..._mat.m:499-521)
        places = [case.force_xy] + (list(case.sensors) if case.sensor_kind == "pzt" else [])
        X = coords
        pzt_el = []
        for xy in places:
            pn, X, _ = pzt.surface_transducer(nodes, X, n, nz, footprint(xy), case.pzt_t)
            pzt_el.append(np.asarray(pn, np.int32))
        n_plate = nodes.shape[0]
        all_nodes = np.vstack([nodes] + pzt_el).astype(np.int32)
        em = np.zeros(all_nodes.shape[0], np.int32)
        if elem_mat is not None:
            em[:n_plate] = elem_mat
        em[n_plate:] = coef.shape[0]
        elem_mat = em
        coef = np.vstack([coef, pzt.solid_coef_row(Q, rho_p)[None, :]])
        fa = pzt.actuator(pzt_el[0], X, n, nz, e, g, case.V0, frame=1)
        dofs = np.nonzero(fa)[0]
        if case.sensor_kind == "pzt":
            outs = []
            for pn in pzt_el[1:]:
                w = pzt.sensor(pn, X, n, nz, e, g, frame=1)
                nzw = np.nonzero(w)[0]
                outs.append((nzw, w[nzw]))
        else:
            outs = [3 * idx_top[nearest_node(coords[idx_top], s)] + 2 for s in case.sensors]
        out.update(nodes=all_nodes, coords=X, coef=coef, elem_mat=elem_mat, ex_dofs=dofs, ex_amp=fa[dofs],
                   outputs=outs, pzt_elems=[np.arange(n_plate, n_plate + len(pzt_el[0]))])
    elif case.transducers == "pzt":
        from . import pzt
        _, e, g, _ = pzt.pzt5a()
        n, nz = case.order + 1, case.order_z + 1
        hx = case.lx / case.nex

        def footprint(xy):
            c = (round(xy[0] / hx) * hx, round(xy[1] / hx) * hx)  # centred on an element corner
            return pzt.top_elements(nodes, coords, c, case.pzt_size / 2, n, nz)

        dofs, amp = pzt.surface_actuator(nodes, coords, n, nz, footprint(case.force_xy), case.pzt_t, e, g, case.V0)
        if case.sensor_kind == "pzt":
            outs = [pzt.surface_sensor(nodes, coords, n, nz, footprint(s), case.pzt_t, e, g) for s in case.sensors]
        else:
            outs = [3 * idx_top[nearest_node(coords[idx_top], s)] + 2 for s in case.sensors]
        out.update(ex_dofs=dofs, ex_amp=amp, outputs=outs)
    else:
        J = idx_top[nearest_node(coords[idx_top], case.force_xy)]
        out.update(ex_dofs=np.array([3 * J + 2]), ex_amp=np.array([case.V0]),
                   outputs=[3 * idx_top[nearest_node(coords[idx_top], s)] + 2 for s in case.sensors])
    return out


def solid_solver(case: SolidCase, prep=None, device=0, kernel=DEFAULT_KERNEL):
    """Configured Solver for a solid case; kernel = (variant, patch elements) or None for
    the library default."""
    prep = prepare_solid(case) if prep is None else prep
    s = Solver(device)
    if kernel is not None:
        s.set_kernel(*kernel)
    s.upload_mesh(SEM_SOLID3, prep["nodes"], prep["coords"], case.order + 1, case.order_z + 1)
    s.set_material(prep["coef"], elem_mat=prep.get("elem_mat"), stiff_scale=case.stiff_scale)
    s.set_damping([case.etad_xy, case.etad_xy, case.etad_z])
    s.set_excitation(prep["ex_dofs"], prep["ex_amp"], prep["st"])
    s.set_outputs(prep["outputs"], case.out_quantity)
    if prep["frame_no"]:
        s.set_frames(prep["frame_no"], "disp_post")
    return s, prep


def timed(fn, *a, **k):
    t0 = time.perf_counter()
    r = fn(*a, **k)
    return r, time.perf_counter() - t0
