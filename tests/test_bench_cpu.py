"""bench.py contract on the CPU: more than one GPU is refused with a message and
status 2, the defaults are the one-GPU headline run, and the CPU-baseline host
record is filled."""
import importlib.util
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench_module():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    out = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_defaults_are_the_one_gpu_headline_run(monkeypatch):
    bench = _bench_module()
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    args = bench.parse()
    assert args.gpus == 1 and args.workload == "tebd"
    assert args.steps == 12 and args.warmup == 0
    assert args.n == 128 and args.chi == 256
    # the other workloads keep their own default of 6 steps
    monkeypatch.setattr(sys, "argv", ["bench.py", "--workload", "batch"])
    assert bench.parse().steps == 6


def test_host_record_of_the_cpu_baseline():
    info = _bench_module().host_info()
    assert set(info) == {"cpu", "logical_cores", "affinity"}
    assert info["logical_cores"] >= 1 and 1 <= info["affinity"] <= info["logical_cores"]
