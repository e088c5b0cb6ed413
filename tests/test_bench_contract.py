"""The bench's command-line contract that needs no GPU: more than one GPU is refused with a
one-line JSON message and exit code 2 before anything touches a device; every step range
the bench and its autotuner run stays inside the signal the case was built with."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def _bench_module():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_tables_and_rates():
    """The committed counter tables are found per workload, form and element kernel; each
    kernel's rate is its counted bytes over its time, and missing tables give None."""
    b = _bench_module()
    t = b.load_traffic("solid100", "fp32_split", (1, 4))
    assert t["bytes_per_step"] == 220530000 and t["source"].startswith("profiles/")
    total = sum(t[k]["read"] + t[k]["write"] for k in ("patch_kernel", "boundary_kernel"))
    assert abs(total - t["bytes_per_step"]) < 1e5
    r = b.kernel_traffic_rates(t, [0.05, 0.02])
    pk = t["patch_kernel"]["read"] + t["patch_kernel"]["write"]
    assert r["patch_kernel"]["GB_per_s"] == pk / 0.05e-3 / 1e9
    assert b.load_traffic("no_such_workload", "fp64", (1, 4)) is None
    # an entry may name its own summary; the others take the table's
    assert b.load_traffic("solid250", "fp32_split", (3, 8))["source"] == "profiles/solid250_staged_kernel_traffic.md"
    assert b.load_traffic("batch256", "fp32_split", (1, 16))["source"] == "profiles/batch256_shell_line_traffic.md"
    assert b.load_traffic("solid250", "fp32_split", (1, 8)) is None  # no table for that kernel
    assert b.kernel_traffic_rates(None, [0.05, 0.02]) is None


class StubSolver:
    """Records the step ranges it is asked to run and refuses, as sem_run does
    (SEM_E_ARG, 'signal shorter than the last step'), any step beyond the signal."""

    def __init__(self, nft, log, n_out=2, dof=6):
        self.nft, self.log, self.n_out, self.dof = nft, log, n_out, dof

    def _check(self, n0, n1):
        from semhip._lib import SemError, SEM_E_ARG
        if n0 < 1 or n1 < n0 or n1 > self.nft:
            raise SemError(SEM_E_ARG, f"signal shorter than the last step ({n0}..{n1} of {self.nft})")
        self.log.append((n0, n1))

    def run(self, dt, n0, n1, precision="fp64", ld=None, want_frames=True, traces_out=None):
        self._check(n0, n1)
        ld = n1 if ld is None else ld
        buf = np.zeros((self.n_out, ld)) if traces_out is None else traces_out
        assert buf.shape == (self.n_out, ld)
        buf[:, n0 - 1:n1] += 1.0  # a step written twice would show as 2
        return dict(traces=buf.T, frames=None, t_frames=None, nn_done=n1, seconds=1e-3 * (n1 - n0 + 1),
                    status="ok")

    def profile(self, dt, n0, nsteps, precision="fp64"):
        self._check(n0, n0 + nsteps - 1)
        return np.zeros(3)

    def set_state(self, *a):
        pass

    def close(self):
        pass


CASES = {"solid250": (30000, 2), "solid100": (20000, 2), "shell40": (5000, 1)}


@pytest.mark.parametrize("config", sorted(CASES))
@pytest.mark.parametrize("steps,warmup", [(20, 5), (2000, 100), (1, 0), (40000, 10)])
def test_every_step_range_fits_the_signal(config, steps, warmup):
    """The driver's `--steps 20 --warmup 5` (and the default, a one-step run, and a timed
    window longer than the workload) on every single-case config: the autotuner's runs,
    the warm-up, timed, rest and profile runs all stay inside the case's signal, the timed
    window has exactly --steps steps and the traces cover every step of the run."""
    from semhip import tune
    b = _bench_module()
    cfg_nft, first = CASES[config]
    a = b.parse(["--config", config, "--steps", str(steps), "--warmup", str(warmup)])
    plan = b.plan_steps(cfg_nft, first, a.warmup, a.steps, a.accuracy_last, a.profile_steps)
    assert plan["timed"][1] - plan["timed"][0] + 1 == steps
    assert plan["nft"] >= cfg_nft
    log = []
    t = tune.autotune(lambda k: StubSolver(plan["nft"], log), 1e-8, first, "fp64", tune.SOLID_CANDIDATES,
                      nn_last=plan["nft"])
    assert t.refused == {} and all(v is not None for v in t.table.values())
    b.make_solver = lambda case, kind, prep, kernel: StubSolver(plan["nft"], log)

    class Case:
        dt = 1e-8
        sensors = ((0.1, 0.1), (0.2, 0.2))

    r = b.run_form(Case(), "solid", {"frame_no": []}, "fp64", plan, (1, 8))
    assert r["traces"].shape[0] == plan["acc_last"] - first + 1
    assert (r["traces"] == 1.0).all()  # every step of first..acc_last ran exactly once
    assert all(n1 <= plan["nft"] for _, n1 in log)


def test_timed_window_in_the_middle_of_the_run():
    """By default the timed window (after its warm-up) starts at the middle of the run, once
    the wave has spread; the run-up before it and the rest after it cover every other step;
    --timed-at 0 times the steps right after the warm-up from the first step."""
    b = _bench_module()
    p = b.plan_steps(30000, 2, 5, 20, 0, 60)
    assert p["timed"] == (15001, 15020) and p["warmup"] == (14996, 15000) and p["lead"] == (2, 14995)
    assert p["rest"] == (15021, 30000)
    q = b.plan_steps(30000, 2, 5, 20, 0, 60, timed_at=0)
    assert q["lead"][1] < q["lead"][0] and q["warmup"] == (2, 6) and q["timed"] == (7, 26)
    r = b.plan_steps(5000, 1, 100, 2000, 0, 0)  # a window longer than the half run
    assert r["timed"] == (2501, 4500) and r["rest"] == (4501, 5000)


@pytest.mark.parametrize("steps,warmup", [(20, 5), (2000, 100), (1, 0)])
def test_batch_step_ranges_fit_the_signal(steps, warmup):
    b = _bench_module()
    p = b.batch_plan(8000, warmup, steps, 0)
    ranges = [p["warmup"], p["timed"], p["rest"], *p["tune"]]
    for n0, n1 in ranges:
        if n1 >= n0:
            assert 1 <= n0 and n1 <= p["nft"]
    assert p["timed"][1] - p["timed"][0] + 1 == steps


def test_autotune_raises_errors_other_than_unsupported():
    """A candidate refused as unsupported is recorded with the library's message; any other
    error (a bad step range here) is raised with the candidate named."""
    from semhip import tune
    from semhip._lib import SemError, SEM_E_UNSUPPORTED

    class Refuses(StubSolver):
        def run(self, *a, **k):
            raise SemError(SEM_E_UNSUPPORTED, "patch needs more LDS than a CU has")

    log = []
    mk = lambda k: Refuses(100, log) if k == (0, 4) else StubSolver(100, log)
    t = tune.autotune(mk, 1e-8, 2, "fp64", tune.SOLID_CANDIDATES, nn_last=100)
    assert t.table[(0, 4)] is None and "LDS" in t.refused[(0, 4)]
    with pytest.raises(SemError) as e:
        tune.autotune(lambda k: StubSolver(50, log), 1e-8, 2, "fp64", tune.SOLID_CANDIDATES)
    assert "element kernel" in str(e.value) and e.value.code != SEM_E_UNSUPPORTED
    assert tune.tune_window(2, 10) == (0, 9)
    assert tune.tune_window(2, 100) == (30, 60)
