"""The bench's command-line contract that needs no GPU: more than one GPU is refused with a
one-line JSON message and exit code 2 before anything touches a device."""
import json
import subprocess
import sys
from pathlib import Path

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
    assert b.kernel_traffic_rates(None, [0.05, 0.02]) is None
