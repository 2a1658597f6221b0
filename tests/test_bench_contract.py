"""bench.py pieces that need no GPU: the --gpus guard and the HBM traffic
figure read from the committed counter summary of the headline workload."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_traffic_from_counter_summary():
    m = _bench()
    t = m.measured_traffic(48, 3)
    assert t is not None
    # within the measured 1.04-1.06 x of the 824.8 MB algorithmic bytes
    alg = m.apply_bytes(9082800, 48 ** 3, 3)
    assert 1.0 <= t / alg <= 1.2, (t, alg)
    assert m.measured_traffic(7, 1) is None


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
