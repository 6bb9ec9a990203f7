"""bench.py on the CPU: it imports, refuses more than one GPU with a message,
and finds the committed HBM-traffic summary only for the workload and tile it
was measured on."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_bench_imports_and_reads_traffic():
    import bench
    t = bench.measured_traffic((512, 256, 256), 16, (8, 4, 4))
    assert t is not None and 5e10 < t < 1e11
    assert bench.measured_traffic((512, 256, 256), 16, (4, 4, 4)) is None
    assert bench.measured_traffic((256, 256, 256), 16, (8, 4, 4)) is None


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
