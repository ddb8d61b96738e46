"""Build libtwdic.so (HIP kernels for gfx950 + C ABI + C++ class layer).

Compiles every source on each call (hipcc cross-compiles without a GPU), in
parallel, into tw_amd/lib/libtwdic.so.  The library is loaded in-tree by
tw_amd.api; nothing is installed.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
CPPSRC = PKG / "cpp"
INCLUDE = ROOT / "include"
OUT_DIR = PKG / "lib"
OBJ_DIR = PKG / "lib" / "obj"
LIB = OUT_DIR / "libtwdic.so"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
ARCH = "gfx950"

# Packed FP32 ops are disabled for all device code (some v_pk_*_f32 forms have
# given wrong results on gfx950 parts); -ffp-contract is left at the HIP default
# (fast) except where the source disables contraction locally.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-fopenmp", f"-I{INCLUDE}", f"-I{CSRC}"]
# host-code sanitizers for tw_host_check_asan: AddressSanitizer and
# UndefinedBehaviorSanitizer on the host side only (the GPU code is built as
# in the library); each -fsanitize= follows -Xarch_host, and the link line
# says -fno-gpu-sanitize
SAN_COMPILE = ["-O1", "-g", "-fno-omit-frame-pointer",
               "-Xarch_host", "-fsanitize=address",
               "-Xarch_host", "-fsanitize=undefined",
               "-Xarch_host", "-fno-sanitize-recover=undefined"]
SAN_LINK = ["-fno-gpu-sanitize", "-fsanitize=address", "-fsanitize=undefined"]
ASAN_DIR = OBJ_DIR / "asan"
HOST_CHECK = OUT_DIR / "tw_host_check_asan"


def _sources():
    hip = sorted(CSRC.glob("*.hip"))
    hip_host = [CSRC / "twdic_api.cpp", CSRC / "batch_api.cpp", CSRC / "serve.cpp"] + sorted(CPPSRC.glob("*.cpp"))
    host = [CSRC / "synth.cpp"]
    return hip, hip_host, host


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(map(str, cmd))}\n{r.stdout}")
    return r.stdout


def build(verbose: bool = False) -> Path:
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    hip, hip_host, host = _sources()
    jobs = []
    for src in hip:
        obj = OBJ_DIR / (src.stem + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)])
    for src in hip_host:
        obj = OBJ_DIR / (src.stem + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in host:
        obj = OBJ_DIR / (src.stem + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    ASAN_DIR.mkdir(parents=True, exist_ok=True)
    san_srcs = hip_host + [CPPSRC / "examples" / "host_check.cpp"]
    for src in san_srcs:
        jobs.append([HIPCC, *HIP_FLAGS, *SAN_COMPILE, "-c", str(src), "-o", str(ASAN_DIR / (src.stem + ".o"))])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [str(OBJ_DIR / (s.stem + ".o")) for s in hip + hip_host + host]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lgomp"])
    os.replace(tmp, LIB)
    # the host-code check with sanitizers: GPU objects as above, host objects
    # instrumented, one executable (the sanitizer runtimes linked in)
    san_objs = [str(OBJ_DIR / (s.stem + ".o")) for s in hip + host] + [str(ASAN_DIR / (s.stem + ".o")) for s in san_srcs]
    _run([HIPCC, f"--offload-arch={ARCH}", *SAN_LINK, "-o", str(HOST_CHECK), *san_objs, "-lgomp"])
    # and TW_Core-style worker drivers, linked against the library in-tree
    exes = [(OUT_DIR / "tw_frame_loop", "frame_loop.cpp"), (OUT_DIR / "tw_icgn_worker", "icgn_worker.cpp"),
            (OUT_DIR / "tw_fftcc_host", "fftcc_host.cpp")]
    for exe, src in exes:
        _run([HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", str(CPPSRC / "examples" / src),
              "-o", str(exe), f"-L{OUT_DIR}", "-ltwdic", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print(f"built {LIB}, {HOST_CHECK} and {', '.join(str(e) for e, _ in exes)}")
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
