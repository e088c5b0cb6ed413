"""Autotuner for the element kernel: times each candidate (variant, patch size) on a few
steps of the real case and precision, and keeps the fastest.

The candidates give the same results to rounding (the GPU tests hold every variant to the
oracle); what differs is how the GLL contractions map onto lanes and LDS, and how many
elements share a workgroup, which trades occupancy against patch-boundary traffic.  A
patch size is fixed when the mesh is uploaded, so each candidate gets its own context.
"""
from __future__ import annotations

from dataclasses import dataclass

from ._lib import SemError, SEM_E_UNSUPPORTED

__all__ = ["SemError", "TuneResult", "autotune", "tune_window", "SOLID_CANDIDATES", "SHELL_CANDIDATES",
           "TUNE_WARMUP", "TUNE_STEPS"]

SOLID_CANDIDATES = ((0, 4), (1, 4), (1, 8), (3, 4), (3, 8))
SHELL_CANDIDATES = ((0, 16), (1, 16), (1, 8), (2, 16))
TUNE_WARMUP = 30
TUNE_STEPS = 60


@dataclass
class TuneResult:
    kernel: tuple
    ms_per_step: float
    table: dict    # candidate -> ms per step (None: refused)
    refused: dict  # candidate -> the library's message for a refused candidate


def tune_window(nn_first, nn_last, warmup=TUNE_WARMUP, steps=TUNE_STEPS):
    """(warmup, steps) of the tuning runs clipped to the steps the case's signal covers:
    the runs use steps nn_first .. nn_first + warmup + steps - 1 <= nn_last.  Short cases
    keep at least one timed step and give up warm-up first."""
    avail = nn_last - nn_first + 1
    if avail < 1:
        raise ValueError(f"no steps to tune on (nn_first {nn_first} > nn_last {nn_last})")
    steps = max(1, min(steps, avail))
    warmup = max(0, min(warmup, avail - steps))
    return warmup, steps


def autotune(make_solver, dt, nn_first, precision, candidates, steps=TUNE_STEPS, warmup=TUNE_WARMUP, nn_last=None):
    """``make_solver(kernel)`` returns a configured Solver whose state is at rest; the
    candidates run steps nn_first.. for ``warmup`` + ``steps`` steps (device time of the
    timed part from the library's own events), clipped to ``nn_last`` (the last step the
    case's signal covers) when given.

    Only a candidate the library refuses as unsupported for this case (variant and patch
    size, or not enough LDS: SEM_E_UNSUPPORTED) is skipped; any other error is raised with
    the candidate named, so that a wrong step range or a device error is not mistaken for
    a refusal."""
    if nn_last is not None:
        warmup, steps = tune_window(nn_first, nn_last, warmup, steps)
    table, refused = {}, {}
    for k in candidates:
        s = None
        try:
            s = make_solver(k)
            if warmup:
                s.run(dt, nn_first, nn_first + warmup - 1, precision, want_frames=False)
            r = s.run(dt, nn_first + warmup, nn_first + warmup + steps - 1, precision, want_frames=False)
            table[k] = 1e3 * r["seconds"] / steps
        except SemError as e:
            if e.code != SEM_E_UNSUPPORTED:
                raise SemError(e.code, f"element kernel {k}: {e}") from e
            table[k] = None
            refused[k] = str(e)
        finally:
            if s is not None:
                s.close()
    ok = {k: v for k, v in table.items() if v is not None}
    if not ok:
        raise SemError(SEM_E_UNSUPPORTED, "no element kernel accepted this case: " +
                       "; ".join(f"{k}: {m}" for k, m in refused.items()))
    best = min(ok, key=ok.get)
    return TuneResult(kernel=best, ms_per_step=ok[best], table=table, refused=refused)
