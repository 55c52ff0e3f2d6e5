"""bench.py's command line, without a GPU: every --workload choice has its
function, and --gpus N > 1 exits with a message (one GPU only)."""
import importlib.util
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_every_workload_has_its_function():
    mod = _bench()
    for name in ("bench_wannier", "bench_velocity", "bench_serving", "bench_batch"):
        assert callable(getattr(mod, name, None)), name
    src = (ROOT / "bench.py").read_text()
    for w in ("haldane", "wannier16", "serving", "batch", "velocity"):
        assert f'"{w}"' in src


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)
