"""Autotuner for the element kernel: times each candidate (variant, patch size) on a few
steps of the real case and precision, and keeps the fastest.

The candidates give the same results to rounding (the GPU tests hold every variant to the
oracle); what differs is how the GLL contractions map onto lanes and LDS, and how many
elements share a workgroup, which trades occupancy against patch-boundary traffic.  A
patch size is fixed when the mesh is uploaded, so each candidate gets its own context.
"""
from __future__ import annotations

from dataclasses import dataclass

from ._lib import SemError

__all__ = ["SemError", "TuneResult", "autotune", "SOLID_CANDIDATES", "SHELL_CANDIDATES"]

SOLID_CANDIDATES = ((0, 4), (1, 4), (1, 8))
SHELL_CANDIDATES = ((0, 16), (1, 16), (1, 8), (2, 16))


@dataclass
class TuneResult:
    kernel: tuple
    ms_per_step: float
    table: dict  # candidate -> ms per step (None: refused, e.g. not enough LDS)


def autotune(make_solver, dt, nn_first, precision, candidates, steps=60, warmup=30):
    """``make_solver(kernel)`` returns a configured Solver whose state is at rest; the
    candidates run steps nn_first.. for ``warmup`` + ``steps`` steps (device time of the
    timed part from the library's own events)."""
    table = {}
    for k in candidates:
        s = None
        try:
            s = make_solver(k)
            s.run(dt, nn_first, nn_first + warmup - 1, precision, want_frames=False)
            r = s.run(dt, nn_first + warmup, nn_first + warmup + steps - 1, precision, want_frames=False)
            table[k] = 1e3 * r["seconds"] / steps
        except SemError:
            table[k] = None
        finally:
            if s is not None:
                s.close()
    ok = {k: v for k, v in table.items() if v is not None}
    if not ok:
        raise SemError(-6, "no element kernel accepted this case")
    best = min(ok, key=ok.get)
    return TuneResult(kernel=best, ms_per_step=ok[best], table=table)
