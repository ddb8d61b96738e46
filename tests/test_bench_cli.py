"""bench.py's contract on the CPU: asked for more than one GPU it exits with
status 2 and says why, before touching a device (the build targets one GPU;
the multi-GPU layout is a design note)."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.parametrize("extra", [[], ["--workload", "batch"], ["--workload", "serving"]])
def test_more_than_one_gpu_is_refused(extra):
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2", *extra], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, (r.returncode, r.stderr[-500:])
    assert "one GPU" in r.stderr
    assert not r.stdout.strip()
