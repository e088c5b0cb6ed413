#!/usr/bin/env python3
"""Benchmark of the spectral-element time step on one MI355X.

Workload (default, --config solid250): 3D laminated plate, 1.0 x 1.0 x 0.004 m,
16-ply [0/45/-45/90]2s CFRP, 250 x 250 x 2 hexahedra of order 5 (6x6x6 GLL),
3 DOF/node (51,645,033 DOF), dt = 2e-8 s, 30000 steps; a 10 mm square, 0.2 mm PZT-5A
actuator at the centre driven by a 75 kHz 5-cycle burst of 10 V; w displacement at 64
surface points on an 8 x 8 grid (plus one 10 mm from the actuator, so that the forms are
compared on a live signal from the first steps on); 20 snapshots, one every 1500 steps.
A "step" is one central-difference step of the whole plate (element forces + assembly +
update + sensors + the snapshots that fall in it).

Each arithmetic form runs the workload's whole step range (steps 2..30000, the 3D
engine's loop); exactly --steps of them, after --warmup untimed steps, are timed.  The
sensor traces and the last snapshot of the whole run are compared between the forms.

Metric: DOF-steps/s of the chosen arithmetic form, the fastest one whose sensor traces
over the whole run stay within 1e-3 relative L2 of the plain FP64 form (the FP64 form
itself is held to the FP64 oracle at 1e-10 by the test suite).  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "ma-shm_amd"))

HBM_PEAK_GBS = 8000.0
PLAIN_SHELL_KERNEL = (1, 8)
FORM_TOL = 1e-3
CONFIGS = ("solid250", "solid100", "shell40", "batch", "adjoint")


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--form", default="auto", choices=["auto", "fp64", "fp32_split"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--profile-steps", type=int, default=60)
    ap.add_argument("--config", default="solid250", choices=CONFIGS)
    ap.add_argument("--accuracy-last", type=int, default=0,
                    help="last step of the run whose traces compare the forms (0: the workload's own "
                         "step count; never before the timed window's last step)")
    ap.add_argument("--timed-at", type=int, default=-1,
                    help="first step of the timed window (-1: the middle of the run, after an untimed run-up; "
                         "0: right after the warm-up from the first step)")
    ap.add_argument("--cases", type=int, default=256, help="batch: number of plates")
    ap.add_argument("--fd", type=int, default=20,
                    help="adjoint: elements of the central-difference gradient check (0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=0,
                    help="adjoint: steps of the run (0: the workload's 15000)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="adjoint: time only the checkpointed sweeps (sem_adjoint) with a seeded synthetic "
                         "misfit derivative, no forward signal runs (kernel traces of the adjoint)")
    ap.add_argument("--kernel", default="auto",
                    help="element kernel 'variant,patch' (e.g. 1,8), or 'auto' to autotune each form")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the all-core CPU leg (0: the CPUs this process may use, at most 16)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only time the CPU baseline of --config and print it as JSON (no GPU)")
    return ap.parse_args(argv)


# ---------------------------------------------------------------------------- step plan
def plan_steps(cfg_nft, first, warmup, steps, accuracy_last, prof_steps, timed_at=-1):
    """Step ranges of one form's runs (1-based, inclusive) and the signal length that
    covers all of them and the autotuner's runs.

    lead    first .. w0-1                      (untimed run-up; empty when w0 = first)
    warmup  w0 .. w0+warmup-1                  (untimed; empty when warmup = 0)
    timed   w0+warmup .. w0+warmup+steps-1
    rest    timed_last+1 .. acc_last           (untimed; acc_last = max(accuracy_last or
                                                cfg_nft, timed_last))
    profile acc_last+1 .. acc_last+prof_steps  (per-kernel events)
    tune    first .. first+TUNE_WARMUP+TUNE_STEPS-1, clipped to the signal by tune_window

    The timed window starts at step timed_at (0: right after the warm-up from the first step;
    -1: the middle of the run, once the wave field has spread over the plate, which is where
    the step costs what it costs over most of the run -- the first steps of a run from rest
    move mostly zeros and run faster); the run-up to it is part of the workload's run.
    """
    from semhip.tune import TUNE_WARMUP, TUNE_STEPS
    if steps < 1 or warmup < 0 or prof_steps < 0:
        raise ValueError("steps >= 1, warmup >= 0 and profile steps >= 0")
    run_last = accuracy_last if accuracy_last > 0 else cfg_nft
    if timed_at < 0:
        timed_at = first + (run_last - first + 1) // 2
    w0 = max(first, timed_at - warmup) if timed_at > 0 else first
    timed_first = w0 + warmup
    timed_last = timed_first + steps - 1
    acc_last = max(accuracy_last if accuracy_last > 0 else cfg_nft, timed_last)
    prof_first = acc_last + 1
    prof_last = acc_last + prof_steps
    nft = max(cfg_nft, prof_last, first + TUNE_WARMUP + TUNE_STEPS - 1)
    return dict(first=first, lead=(first, w0 - 1), warmup=(w0, timed_first - 1), timed=(timed_first, timed_last),
                rest=(timed_last + 1, acc_last), profile=(prof_first, prof_last), acc_last=acc_last, nft=nft)


def batch_plan(cfg_nft, warmup, steps, accuracy_last, timed_at=-1):
    """The batch's step ranges (shells, loop nn = 1..nft): tuning 1..10 kernel by kernel and
    11..30 as one graph chunk, then per form the run-up, warm-up, timed window and the rest
    of the run up to acc_last; the signal covers all of them."""
    p = plan_steps(cfg_nft, 1, warmup, steps, accuracy_last, 0, timed_at)
    p["tune"] = ((1, 10), (11, 30))
    p["nft"] = max(p["nft"], 30)
    return p


# ---------------------------------------------------------------------------- workloads
def build_case(name):
    """The workload plus one sensor next to the source, so that the form comparison sees
    the wave from the first steps on as well as the far sensors."""
    from semhip import engine as E
    if name == "solid100":
        c = E.solid_plate_100x100_pzt()
        c.sensors = tuple(c.sensors) + ((c.force_xy[0] + 0.02, c.force_xy[1]),)
        return c, "solid"
    if name == "solid250":
        c = E.solid_plate_250x250_pzt()
        c.sensors = tuple(c.sensors) + ((c.force_xy[0] + 0.01, c.force_xy[1]),)
        return c, "solid"
    c = E.shell_plate_40x40()
    c.sensors = tuple(c.sensors) + ((c.force_xy[0] + 0.0125, c.force_xy[1]),)
    return c, "shell"


WORKLOAD = {
    "solid250": "3D solid 250x250x2 order-5 laminated plate [0/45/-45/90]2s, PZT-5A actuator 10 V 75 kHz, "
                "64 surface sensors (+1 near the actuator), 20 snapshots over 30000 steps",
    "solid100": "3D solid 100x100x2 order-4 laminated plate [0/90/0/90]s, PZT-5A actuator 10 V 100 kHz + "
                "8 PZT voltage sensors (+1 near the actuator), 20000 steps",
    "shell40": "flat shell 40x40 order-5 laminated plate, point force 50 kHz, 4 w-velocity sensors (+1), "
               "10 snapshots over 5000 steps",
}


def copy_calibration(n_bytes=1 << 30, reps=5):
    """A plain device copy of known size (n_bytes read + n_bytes written per launch), run
    under the same counters as the step kernels to scale FETCH_SIZE / WRITE_SIZE."""
    import torch
    src = torch.ones(n_bytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n_bytes, "launches": reps}), flush=True)


def load_traffic(config, form, kernel):
    """HBM traffic per step of this workload's step kernels for one form and element kernel
    (variant, patch size), measured by rocprofv3 PMC runs committed with their commands under
    profiles/, or None."""
    f = ROOT / "profiles" / f"traffic_{config}.json"
    if not f.exists():
        return None
    try:
        tbl = json.loads(f.read_text())
        t = tbl.get(form, {}).get(f"{kernel[0]},{kernel[1]}")
        return dict(t, source=t.get("source", tbl["source"])) if t else None
    except (ValueError, OSError, KeyError, TypeError):
        return None


def load_json_table(name):
    """A committed counter table under profiles/ (None when absent or unreadable)."""
    try:
        return json.loads((ROOT / "profiles" / name).read_text())
    except (ValueError, OSError):
        return None


def kernel_traffic_rates(traffic, kernel_ms):
    """Each step kernel's measured HBM bytes (read + write, from the committed counter runs)
    over its own mean duration (HIP events around each launch)."""
    if not traffic or "patch_kernel" not in traffic or kernel_ms is None or len(kernel_ms) < 2:
        return None
    out = {}
    for name, ms in (("patch_kernel", kernel_ms[0]), ("boundary_kernel", kernel_ms[1])):
        b = traffic[name]["read"] + traffic[name]["write"]
        out[name] = {"ms": float(ms), "bytes": b, "GB_per_s": (b / (ms * 1e-3) / 1e9) if ms > 0 else None}
    return out


def make_solver(case, kind, prep, kernel):
    from semhip import engine as E
    mk = E.solid_solver if kind == "solid" else E.flat_shell_solver
    return mk(case, prep, kernel=kernel)[0]


def run_form(case, kind, prep, form, plan, kernel):
    """One form over the whole plan: warm-up, the timed window, the rest of the run, then
    the per-kernel profile.  Returns the traces of every step, the last snapshot and the
    timed window's device seconds (HIP events on the context's stream) and wall time."""
    s = make_solver(case, kind, prep, kernel)
    ld = plan["acc_last"]
    last_frame, last_frame_nn = None, None
    # one trace buffer for all of the form's runs, each run writing its own rows; allocated and
    # touched before the timed run, so that the timed region holds no first-touch page faults
    tbuf = np.zeros((len(case.sensors), ld))
    tbuf.fill(0.0)

    def take(r, n0, n1):
        # frames: the snapshots whose step is in n0..n1
        nonlocal last_frame, last_frame_nn
        got = [i for i, fn in enumerate(prep["frame_no"]) if n0 <= fn <= n1]
        if got and r["frames"] is not None:
            last_frame, last_frame_nn = r["frames"][got[-1]].copy(), got[-1]

    def frames_in(n0, n1):
        # the device samples every snapshot of a run either way; the host buffer for them
        # (frames x dof doubles) is only allocated when one falls in the run
        return any(n0 <= fn <= n1 for fn in prep["frame_no"])

    for w0, w1 in (plan["lead"], plan["warmup"]):
        if w1 >= w0:
            take(s.run(case.dt, w0, w1, form, ld=ld, want_frames=frames_in(w0, w1), traces_out=tbuf), w0, w1)
    t0 = time.perf_counter()
    r = s.run(case.dt, *plan["timed"], form, ld=ld, want_frames=frames_in(*plan["timed"]), traces_out=tbuf)
    wall = time.perf_counter() - t0
    take(r, *plan["timed"])
    seconds = r["seconds"]
    a0, a1 = plan["rest"]
    if a1 >= a0:
        take(s.run(case.dt, a0, a1, form, ld=ld, want_frames=frames_in(a0, a1), traces_out=tbuf), a0, a1)
    p0, p1 = plan["profile"]
    ms = s.profile(case.dt, p0, p1 - p0 + 1, form) if p1 >= p0 else np.zeros(3)
    s.close()
    return dict(traces=tbuf.T[plan["first"] - 1:], seconds=seconds, wall=wall, kernel_ms=ms, last_frame=last_frame,
                last_frame_index=last_frame_nn)


# ---------------------------------------------------------------------------- CPU baseline
def cpu_threads(requested=0):
    """Threads of the all-core leg: the CPUs this process may run on, at most 16 (the host
    share of one GPU on the benchmark machines)."""
    avail = len(os.sched_getaffinity(0))
    return max(1, min(avail, 16) if requested <= 0 else requested)


def cpu_baseline(config, kind, n_threads):
    """The oracle (NumPy/SciPy restatement of the reference's vectorised mode='cpu' loop)
    on a bounded sample of the same workload, timed twice on the host cores: one thread
    (the paper's single-core method) and all cores (the element stage over contiguous
    element ranges, one worker process per core; assembly and update as in the reference).

    solid250: a 20 x 20 x 2-element part of the same plate (same 4 mm elements, order 5,
              16-ply laminate, dt), 60 steps per leg after one warm-up step;
    solid100: a 40 x 40 x 2-element part of the 100 x 100 x 2 plate (order 4), 120 steps;
    shell40:  the whole 40 x 40 plate, 300 steps."""
    from threadpoolctl import threadpool_limits
    from oracle import mesh as om, solid as oso, shell as osh
    from oracle.common import ChunkedForces
    if kind == "solid":
        if config == "solid250":
            n_el, order, h, nsteps = 20, 5, 4e-3, 60
            lh = [0.25e-3] * 16
            ang = [0, 45, -45, 90, 0, 45, -45, 90, 90, -45, 45, 0, 90, -45, 45, 0]
            side = 20 * 1.0 / 250
        else:
            n_el, order, h, nsteps = 40, 4, 2e-3, 120
            lh = [0.25e-3] * 8
            ang = [0, 90, 0, 90, 90, 0, 90, 0]
            side = 40 * 0.5 / 100
        q, p = om.rect_quad_mesh(side, side, n_el, n_el)
        n2, c2 = om.quad2spectral(q, p, order)
        n3, c3 = om.extrude_mesh(n2, c2, order + 1, order + 1, order + 1, [h / 2, h / 2], 2)
        Q = oso.lamina_q3d(130e9, 10e9, 10e9, 5e9, 5e9, 3.5e9, 0.3, 0.3, 10 / 7 - 1)
        lam = oso.laminate_solid_smeared(lh, ang, 1570, **Q)
        model = oso.SolidModel(n3, c3, order + 1, order + 1, [lam[k] for k in oso.SOLID_COEFS])
        sample = (f"{n_el}x{n_el}x2 hexahedra of order {order} (same element size, laminate and dt as the "
                  f"workload), {nsteps} steps per leg")
        dt, first = 2e-8, 2
    else:
        q, p = om.rect_quad_mesh(0.5, 0.5, 40, 40)
        n2, c2 = om.quad2spectral(q, p, 5)
        Q = osh.lamina_q(130e9, 10e9, 5e9, 5e9, 3.5e9, 0.3)
        lam = osh.laminate_shell([0.25e-3] * 8, [0, 90, 0, 90, 90, 0, 90, 0], 1570, **Q)
        model = osh.ShellModel(n2, c2, 5, np.array([[lam[k] for k in osh.SHELL_COEFS]]))
        sample = "40x40 shells of order 5 (the whole plate), 300 steps per leg"
        nsteps, dt, first = 300, 1e-7, 1
    st = np.ones(nsteps + first + 2)
    f = np.zeros(model.dof)
    f[model.dof // 2 - (model.dof // 2) % model.nf + model.nf - 1] = 1.0
    legs = {}
    with threadpool_limits(limits=1):
        for label, nt in (("one_thread", 1), ("all_cores", n_threads)):
            model.chunked = ChunkedForces(model, nt) if nt > 1 else None
            model.run(dt, st, f, first, first)  # warm-up
            t0 = time.perf_counter()
            model.run(dt, st, f, first, first + nsteps - 1)
            sec = time.perf_counter() - t0
            if model.chunked is not None:
                model.chunked.close()
            model.chunked = None
            legs[label] = dict(value=model.dof * nsteps / sec, cores=nt, seconds=sec)
    host = host_cpu()
    return dict(value=legs["all_cores"]["value"], unit="DOF-steps/s", cores=legs["all_cores"]["cores"], kind="port",
                sample=f"{sample}; reference-algorithm CPU restatement (oracle/); host {host}",
                one_thread=legs["one_thread"], all_cores=legs["all_cores"], dof=model.dof)


def cpu_baseline_child(config, n_threads):
    """The CPU baseline in a child Python process, which has not touched the GPU: its
    all-core leg forks worker processes, which a process holding a HIP context should not."""
    import subprocess
    cmd = [sys.executable, str(Path(__file__).resolve()), "--cpu-only", "--config", config,
           "--cpu-threads", str(n_threads)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def host_cpu():
    """CPU model and the number of CPUs this process may run on."""
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{model}, {len(os.sched_getaffinity(0))} CPUs available"


def rel_l2(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


# ---------------------------------------------------------------------------- batch
def bench_batch(a):
    """Batched damage sweep (256 flat-shell plates of 60..120 elements per side, order 5):
    the FP64 and FP32-element forms of the whole batch over its whole step range (8000
    steps; --steps of them timed), each case's FP32 traces checked against its FP64 solve,
    and the plain FP64 single-plate run (40 x 40 plate) timed in the same process."""
    from semhip import batch as B, tune
    from semhip import engine as E
    from semhip._lib import SEM_E_UNSUPPORTED
    cfg = B.damage_sweep(n_cases=a.cases)
    plan = batch_plan(cfg.nft, a.warmup, a.steps, a.accuracy_last, a.timed_at)
    cfg.nft = plan["nft"]
    t0 = time.perf_counter()
    prep = B.prepare_batch(cfg)
    print(f"[bench] batch: {len(cfg.cases)} plates, dof={prep['dof']} elements={prep['elements']} "
          f"(set-up {time.perf_counter() - t0:.1f} s), plan {plan}", file=sys.stderr, flush=True)
    solvers, table = {}, {f: {} for f in ("fp64", "fp32_split")}
    refused = {}
    cands = tune.SHELL_CANDIDATES if a.kernel == "auto" else (tuple(int(x) for x in a.kernel.split(",")),)
    for k in cands:
        t1 = time.perf_counter()
        try:
            solvers[k] = B.batch_solver(cfg, prep, kernel=k)[0]
        except tune.SemError as e:
            if e.code != SEM_E_UNSUPPORTED:
                raise
            refused[k] = str(e)
            continue
        print(f"[bench] batch upload for kernel {k}: {time.perf_counter() - t1:.1f} s", file=sys.stderr, flush=True)
        for form in table:
            try:
                solvers[k].set_state(None, None)
                # the first range runs kernel by kernel (shorter than a graph chunk) so that
                # counter runs see it; the second is one graph chunk
                (u0, u1), (v0, v1) = plan["tune"]
                solvers[k].run(cfg.dt, u0, u1, form, want_frames=False)
                r = solvers[k].run(cfg.dt, v0, v1, form, want_frames=False)
                table[form][k] = 1e3 * r["seconds"] / (v1 - v0 + 1)
            except tune.SemError as e:
                if e.code != SEM_E_UNSUPPORTED:
                    raise tune.SemError(e.code, f"element kernel {k}: {e}") from e
                table[form][k] = None
                refused[k] = str(e)
    res = {}
    for form in ("fp64", "fp32_split"):
        ok_k = {k: v for k, v in table[form].items() if v is not None}
        if not ok_k:
            raise SystemExit(f"no element kernel ran the {form} batch: {refused}")
        kern = min(ok_k, key=ok_k.get)
        tab = {",".join(map(str, k)): v for k, v in table[form].items()}
        print(f"[bench] batch {form}: autotune ms/step {tab} -> {kern}", file=sys.stderr, flush=True)
        s = solvers[kern]
        s.set_state(None, None)
        ld = plan["acc_last"]
        tr = None
        for w0, w1 in (plan["lead"], plan["warmup"]):
            if w1 >= w0:
                t = s.run(cfg.dt, w0, w1, form, ld=ld, want_frames=False)["traces"]
                tr = t if tr is None else tr + t
        t1 = time.perf_counter()
        r = s.run(cfg.dt, *plan["timed"], form, ld=ld, want_frames=False)
        wall = time.perf_counter() - t1
        tr = r["traces"] if tr is None else tr + r["traces"]
        a0, a1 = plan["rest"]
        if a1 >= a0:
            tr = tr + s.run(cfg.dt, a0, a1, form, ld=ld, want_frames=False)["traces"]
        res[form] = dict(seconds=r["seconds"], wall=wall, traces=tr, kernel=kern, tune=tab)
        print(f"[bench] batch {form}: {r['seconds']:.3f} s device, {wall:.3f} s wall (timed window)",
              file=sys.stderr, flush=True)
    for s in solvers.values():
        s.close()
    if not all(np.isfinite(res[f]["traces"]).all() for f in res):
        raise SystemExit("batch run diverged (non-finite sensor traces)")
    ref = B.split(res["fp64"]["traces"], cfg)
    per_case = [rel_l2(x, y) for x, y in zip(B.split(res["fp32_split"]["traces"], cfg), ref)]
    # plain FP64 single plate (config-1 size) for the ratio, on the line-wise kernel with
    # 8-element patches (the autotuner's pick for this plate in both forms)
    c1 = E.shell_plate_40x40(nft=2200)
    s, _ = E.flat_shell_solver(c1, kernel=PLAIN_SHELL_KERNEL)
    s.run(c1.dt, 1, 100, "fp64")
    r1 = s.run(c1.dt, 101, 2100, "fp64", want_frames=False)
    s.close()
    plain = c1.dof * 2000 / r1["seconds"]
    ok = [f for f in res if f == "fp64" or max(per_case) <= FORM_TOL]
    chosen = min(ok, key=lambda f: res[f]["wall"])
    value = prep["dof"] * a.steps / res[chosen]["wall"]
    b_step = 32 * prep["dof"] + 4 * prep["elements"] * (cfg.order + 1) ** 2
    traffic = load_traffic(f"batch{len(cfg.cases)}", chosen, res[chosen]["kernel"])
    same_batch_fp64 = prep["dof"] * a.steps / res["fp64"]["wall"]
    out = {
        "metric": "DOF-steps/s on 1 MI355X, batched damage sweep, chosen vs plain FP64 form at <=1e-3 rel. L2 error",
        "value": value, "unit": "DOF-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": res[chosen]["wall"] / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if chosen == "fp64" else "f32-element/f64-state", "data": "synthetic",
        "config": {"workload": f"{len(cfg.cases)} flat-shell plates 0.5x0.5 m, 60..120 order-5 elements per side, "
                               "elliptical stiffness-scaled region each, centre 50 kHz burst",
                   "seed": cfg.seed, "dof": prep["dof"], "elements": prep["elements"], "form": chosen,
                   "timed_steps": list(plan["timed"]), "untimed_run_up": list(plan["lead"]),
                   "steps_compared": plan["acc_last"], "parallelism": "single GPU"},
        "forms": {f: {"dof_steps_per_s": prep["dof"] * a.steps / res[f]["wall"],
                      "device_seconds": res[f]["seconds"], "element_kernel": list(res[f]["kernel"]),
                      "autotune_ms_per_step": res[f]["tune"]} for f in res},
        "per_case_rel_l2_fp32_vs_fp64": {"max": max(per_case), "median": float(np.median(per_case)),
                                         "steps": plan["acc_last"]},
        "batched_vs_same_batch_fp64": value / same_batch_fp64,
        "mfma": load_json_table("mfma_shell_kernel.json"),
        "plain_fp64_single_plate_dof_steps_per_s": plain,
        "plain_fp64_single_plate_element_kernel": list(PLAIN_SHELL_KERNEL),
        "batched_vs_plain_single_plate": value / plain,
        "roofline": {"bound": "hbm", "achieved": b_step / (res[chosen]["seconds"] / a.steps) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": b_step / (res[chosen]["seconds"] / a.steps) / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_step"] if traffic else None,
                     "element_kernel": list(res[chosen]["kernel"]),
                     "algorithmic_bytes_per_step": b_step,
                     "traffic_source": traffic["source"] if traffic else None},
    }
    print(json.dumps(out), flush=True)


def bench_adjoint(a):
    """Adjoint inversion (120 x 120 order-5 shell, 15000 steps): misfit gradient w.r.t. the
    stiffness and density fields by the discrete adjoint in FP64 and in the FP32 form, and
    central differences (FP64) on --fd random elements (20 by default)."""
    from semhip import inversion as I
    case = I.inversion_120x120(**({"nft": a.adjoint_steps} if a.adjoint_steps > 0 else {}))
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if a.adjoint_only:
        r = I.adjoint_sweeps(case, log=log)
        out = {"metric": "DOF-steps/s of the discrete adjoint sweeps (checkpointed forward + reverse) on 1 MI355X",
               "value": r["dof"] * case.nft / r["fp32_split"]["adjoint_seconds"], "unit": "DOF-steps/s",
               "n_gpus": 1, "steps": case.nft, "warmup": 0, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32-element/f64-state", "data": "synthetic",
               "config": {"workload": "adjoint sweeps alone, 120x120 order-5 flat shell, 9 sensors, seeded "
                                      "synthetic misfit derivative", "dof": r["dof"], "elements": r["elements"],
                          "parallelism": "single GPU"},
               "adjoint": r}
        print(json.dumps(out), flush=True)
        return
    r = I.gradient_check(case, n_fd=a.fd, log=log)
    f32 = r["fp32_split"]
    out = {
        "metric": "DOF-steps/s of the discrete adjoint (checkpointed forward + reverse sweep) on 1 MI355X",
        "value": r["dof"] * case.nft / f32["adjoint_wall"], "unit": "DOF-steps/s", "n_gpus": 1, "steps": case.nft,
        "warmup": 0, "ms_per_step": f32["adjoint_wall"] / case.nft * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-element/f64-state", "data": "synthetic",
        "config": {"workload": "adjoint inversion, 120x120 order-5 flat shell, 9 sensors, Gaussian-blob stiffness "
                               "field", "dof": r["dof"], "elements": r["elements"], "seed": case.seed,
                   "parallelism": "single GPU"},
        "adjoint": r,
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------- main
def main(argv=None):
    a = parse(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "designed, not built)"}))
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    if a.cpu_only:
        kind = "shell" if a.config == "shell40" else "solid"
        print(json.dumps(cpu_baseline(a.config, kind, cpu_threads(a.cpu_threads))), flush=True)
        return
    if a.config == "batch":
        bench_batch(a)
        return
    if a.config == "adjoint":
        bench_adjoint(a)
        return
    for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ.setdefault(v, "1")
    from semhip import engine as E, tune
    t0 = time.perf_counter()
    case, kind = build_case(a.config)
    first = 2 if kind == "solid" else 1
    plan = plan_steps(case.nft, first, a.warmup, a.steps, a.accuracy_last, a.profile_steps, a.timed_at)
    prep = E.prepare_solid(case) if kind == "solid" else E.prepare_flat_shell(case)
    # the signal of the workload's own nft (its frame schedule stays), extended by zeros
    # (the burst has ended) to cover the profile steps after the run
    prep["st"] = np.concatenate([prep["st"], np.zeros(plan["nft"] - prep["st"].size)])
    n_el, n_loc = prep["nodes"].shape
    dof = (3 if kind == "solid" else 5) * prep["coords"].shape[0]
    print(f"[bench] {a.config}: dof={dof} elements={n_el} steps={a.steps} warmup={a.warmup} plan={plan} "
          f"(set-up {time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    forms = ["fp64", "fp32_split"] if a.form in ("auto", "fp32_split") else ["fp64"]
    res, tuned = {}, {}
    for form in forms:
        if a.kernel == "auto":
            cands = tune.SOLID_CANDIDATES if kind == "solid" else tune.SHELL_CANDIDATES
            t = tune.autotune(lambda k: make_solver(case, kind, prep, k), case.dt, first, form, cands,
                              nn_last=plan["nft"])
            kern = t.kernel
            tuned[form] = {",".join(map(str, k)): v for k, v in t.table.items()}
            print(f"[bench] {form}: autotune ms/step {tuned[form]} -> {kern}", file=sys.stderr, flush=True)
        else:
            kern = tuple(int(x) for x in a.kernel.split(","))
        t1 = time.perf_counter()
        res[form] = run_form(case, kind, prep, form, plan, kern)
        res[form]["kernel"] = kern
        r = res[form]
        print(f"[bench] {form}: timed {r['seconds']:.4f} s device, {r['wall']:.4f} s wall; whole run "
              f"{time.perf_counter() - t1:.1f} s; kernels ms {np.round(r['kernel_ms'], 5).tolist()}",
              file=sys.stderr, flush=True)
    ref = res["fp64"]
    if not np.isfinite(ref["traces"]).all():
        raise SystemExit("FP64 run diverged (non-finite sensor traces)")
    errs = {f: rel_l2(res[f]["traces"], ref["traces"]) for f in res}
    ferrs = {f: (rel_l2(res[f]["last_frame"], ref["last_frame"]) if ref["last_frame"] is not None else None)
             for f in res}
    ok = [f for f in res if errs[f] <= FORM_TOL]
    chosen = min(ok, key=lambda f: res[f]["wall"]) if a.form == "auto" else a.form
    r = res[chosen]
    value = dof * a.steps / r["wall"]
    # algorithmic bytes per step: read U^n, U^{n-1}, write U^{n+1}, read one FP64 inverse-mass
    # coefficient (8 B each per DOF), plus 4 B per connectivity entry
    b_step = 32 * dof + 4 * n_el * n_loc
    step_s = r["seconds"] / a.steps
    achieved = b_step / step_s / 1e9 if step_s > 0 else None
    traffic = load_traffic(a.config, chosen, r["kernel"])
    out = {
        "metric": "DOF-steps/s on 1 MI355X, chosen vs plain FP64 form at <=1e-3 rel. L2 error",
        "value": value,
        "unit": "DOF-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["wall"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if chosen == "fp64" else "f32-element/f64-state",
        "data": "synthetic",
        "config": {"workload": WORKLOAD[a.config], "sensors": len(case.sensors), "snapshots": len(prep["frame_no"]),
                   "steps_timed": a.steps, "timed_steps": list(plan["timed"]), "untimed_run_up": list(plan["lead"]),
                   "steps_run": [first, plan["acc_last"]],
                   "dof": dof, "elements": n_el, "dt": case.dt, "form": chosen, "parallelism": "single GPU"},
        "forms": {f: {"dof_steps_per_s": dof * a.steps / res[f]["wall"],
                      "roofline_frac": b_step * a.steps / res[f]["seconds"] / 1e9 / HBM_PEAK_GBS,
                      "rel_l2_vs_fp64_traces": errs[f], "rel_l2_vs_fp64_last_snapshot": ferrs[f],
                      "steps_compared": plan["acc_last"] - first + 1,
                      "device_seconds": res[f]["seconds"], "kernel_ms": [float(x) for x in res[f]["kernel_ms"]],
                      "element_kernel": list(res[f]["kernel"]), "autotune_ms_per_step": tuned.get(f)}
                  for f in res},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic["bytes_per_step"] if traffic else None,
                     "element_kernel": list(r["kernel"]),
                     "kernel": "one time step (patch kernel + boundary-node kernel, plus the per-chunk sensor "
                               "kernel amortised); per-kernel ms in forms",
                     "algorithmic_bytes_per_step": b_step,
                     "traffic_source": traffic["source"] if traffic else None,
                     "measured_traffic_rate": kernel_traffic_rates(traffic, r["kernel_ms"])},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline_child(a.config, cpu_threads(a.cpu_threads))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
