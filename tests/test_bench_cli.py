"""bench.py's command line on the CPU: more than one GPU is refused with a
message and a non-zero status before anything touches a device, and the
flags parse to the documented defaults."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120, cwd=ROOT)
    assert res.returncode == 2
    assert "one GPU only" in res.stderr
    assert res.stdout == ""


def test_flag_defaults(monkeypatch):
    sys.path.insert(0, str(ROOT))
    try:
        import bench
    finally:
        sys.path.remove(str(ROOT))
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    a = bench.parse()
    assert (a.gpus, a.steps, a.warmup, a.samples_log2, a.rpoints) == (1, 100, 2, 20, 255)
    assert a.large_maxy == 0.0 and a.large_precision == "both"
    assert a.batch_runs == 64 and a.dis_points == 32 and not a.kernel_probe
    assert not a.no_config1 and not a.no_cpu_baseline
