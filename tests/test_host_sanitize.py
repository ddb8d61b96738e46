"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tw_amd/lib/tw_host_check_asan (built by tw_amd/build.py) links the library's
host code (C ABI, batch and serving layers) compiled with both sanitizers,
UBSan aborting on its first finding, and the GPU objects as in the library.
Its driver (tw_amd/cpp/examples/host_check.cpp) runs seeded random
configurations through validation and creation, the CSV writer and synthetic
pairs; with a GPU it also runs a context through every entry point, a batch
in both launch forms against the context bit for bit, and a burst of
requests through a server.  Leak detection is off: the HIP runtime keeps its
allocations to the end of the process."""
import os
import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parents[1] / "tw_amd" / "lib" / "tw_host_check_asan"


def _run(timeout):
    if not EXE.exists():
        pytest.fail(f"{EXE} is missing: run tw_amd/build.py")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(EXE)], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout
    return r.stdout


def test_host_code_sanitized_without_device():
    """Part A everywhere (with a GPU, part B runs too, under the gpu test)."""
    out = _run(600)
    assert "host_check:" in out


@pytest.mark.gpu
def test_host_code_sanitized_on_device():
    out = _run(600)
    assert "parts A and B" in out
