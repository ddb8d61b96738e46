"""CPU checks of the product library (no GPU needed): the C ABI loads and
exports every function its headers declare, the drop-in C++ headers are
self-contained, context creation fails cleanly without a GPU, and the
synthetic speckle generator is deterministic with exact displacement fields."""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from tw_amd import api

ROOT = Path(__file__).resolve().parents[1]


def _declared(header: Path):
    text = header.read_text()
    return sorted(set(re.findall(r"^\s*(?:twdic_status|int|void|const char\*|void\*)\s+\**(twdic_\w+)\s*\(",
                                 text, re.M)))


@pytest.mark.parametrize("header", ["twdic.h", "twdic_synth.h", "twdic_batch.h", "twdic_serve.h"])
def test_library_exports_every_declared_symbol(header, lib_built):
    names = _declared(ROOT / "include" / header)
    assert len(names) >= 2
    nm = subprocess.run(["nm", "-D", "--defined-only", str(api.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(\w+)$", nm, re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib_built, n)          # resolvable through ctypes too


def test_ctypes_signatures_cover_the_abi():
    names = set()
    for h in ("twdic.h", "twdic_synth.h", "twdic_batch.h", "twdic_serve.h"):
        names |= set(_declared(ROOT / "include" / h))
    assert names == set(api.exported_symbols())


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="needs hipcc for the HIP headers")
def test_dropin_headers_are_self_contained(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text("""
#include "tw/TW_paDIC_cuFFTCC2D.h"
#include "tw/TW_paDIC_cuICGN2D.h"
struct FakeMat { int rows = 4, cols = 5; size_t step = 5; unsigned char* data = nullptr; };
void use() {
    FakeMat m;
    TW::MatView v = m;                                   // cv::Mat-shaped objects convert implicitly
    (void)v;
    TW::real_t** fU = nullptr;
    TW::hcreateptr(fU, 3, 4);
    fU[2][3] = 1.f;
    TW::hdestroyptr(fU);
    TW::paDIC::GPUHandle_ICGN h;
    (void)h.m_d_f4InterpolationLUT;
}
""")
    r = subprocess.run(["hipcc", "-std=c++17", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_create_without_gpu_fails_cleanly(lib_built):
    import ctypes as C
    ndev = C.c_int(0)
    try:
        hip = C.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount(C.byref(ndev))
    except OSError:
        pass
    cfg = api.make_config(64, 64, 1, 1, 62, 62, 5, 5, 5, 5, 2, 2, 16)
    h = C.c_void_p()
    st = lib_built.twdic_create(C.byref(cfg), 0, C.byref(h))
    if ndev.value == 0:
        assert st == api.ENODEV and not h.value
    # invalid geometry is rejected before any device work
    bad = api.make_config(64, 64, 1, 1, 62, 62, 40, 40, 5, 5, 2, 2, 16)
    assert lib_built.twdic_create(C.byref(bad), 0, C.byref(h)) == api.EINVAL
    bad = api.make_config(64, 64, 1, 1, 62, 62, 5, 5, 5, 5, 2, 2, 33)   # odd window
    assert lib_built.twdic_create(C.byref(bad), 0, C.byref(h)) == api.EINVAL
    # beyond the 32-bit image addressing: 65536 px wide, or 2^31 bytes
    bad = api.make_config(65536, 64, 1, 1, 62, 62, 5, 5, 5, 5, 2, 2, 16)
    assert lib_built.twdic_create(C.byref(bad), 0, C.byref(h)) == api.EINVAL
    bad = api.make_config(46341, 46341, 1, 1, 62, 62, 5, 5, 5, 5, 2, 2, 16)
    assert lib_built.twdic_create(C.byref(bad), 0, C.byref(h)) == api.EINVAL


def test_num_poi_matches_oracle(lib_built):
    rng = np.random.default_rng(4)
    for _ in range(200):
        rw, rh = rng.integers(10, 3000, 2)
        sx, gx, mx = rng.integers(1, 40), rng.integers(1, 20), rng.integers(0, 40)
        cfg = api.make_config(int(rw) + 2, int(rh) + 2, 1, 1, int(rw), int(rh), int(sx), int(sx), int(gx), int(gx),
                              int(mx), int(mx), 32)
        nx, ny, ok = oracle.num_poi(int(rw), int(rh), int(sx), int(sx), int(gx), int(gx), int(mx), int(mx))
        assert api.num_poi(cfg) == (nx, ny)


def test_synth_is_deterministic_and_exact(lib_built):
    d = api.affine(1.5, -0.5, 0.01, -0.02, 0.005, 50.0, 60.0)
    a = api.synth_pair(160, 120, 300, 99, d)
    b = api.synth_pair(160, 120, 300, 99, d)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = api.synth_pair(160, 120, 300, 100, d)
    assert not np.array_equal(a[0], c[0])
    # affine displacement closed form, symmetric shear
    o = api.displacement(d, 70.0, 40.0)
    assert np.allclose(o, [1.5 + 0.01 * 20 + 0.005 * -20, 0.01, 0.005, -0.5 + 0.005 * 20 - 0.02 * -20, 0.005, -0.02])
    s = api.sinusoid(0.5, 256.0, 0.3, 192.0)
    o = api.displacement(s, 64.0, 48.0)
    assert np.allclose(o, [0.5, 0.0, 0.0, 0.0, 0.0, -0.3 * 2 * np.pi / 192], atol=1e-12)
    # a noise-free zero deformation reproduces the reference exactly
    r0, t0 = api.synth_pair(96, 80, 100, 5, api.affine(0, 0, xc=48, yc=40), noise_sigma=0.0)
    assert np.array_equal(r0, t0)


def test_synth_speckle_statistics(lib_built):
    ref, tar = api.synth_pair(512, 512, 4000, 1001, api.resample(2.35, -1.72))
    assert 60 < ref.mean() < 160 and ref.std() > 40
    # translation recovered by phase correlation of the whole images
    F = np.fft.fft2(ref.astype(float) - ref.mean())
    G = np.fft.fft2(tar.astype(float) - tar.mean())
    c = np.real(np.fft.ifft2(np.conj(F) * G))
    k = np.unravel_index(np.argmax(c), c.shape)
    assert (k[1], k[0] - 512) == (2, -2)


def test_poi_count_bounds_check_on_the_host():
    """tw_amd.api refuses a POI count its per-POI buffers cannot hold, before
    any call into the library (the C ABI sees plain pointers only)."""
    from types import SimpleNamespace as B

    from tw_amd import api
    n = 10
    ok = dict(poi=B(nbytes=8 * n), u=B(nbytes=4 * n), v=B(nbytes=4 * n), p6=B(nbytes=24 * n),
              iters=B(nbytes=4 * n), flags=B(nbytes=n))
    api._fits(n, **ok)
    api._fits(0, **ok)
    api._fits(n, poi=ok["poi"], u=ok["u"], v=ok["v"], p6=None, flags=None)   # optional outputs absent
    for k in ok:
        short = dict(ok, **{k: B(nbytes=ok[k].nbytes - 1)})
        with pytest.raises(ValueError, match=k):
            api._fits(n, **short)
    with pytest.raises(ValueError):
        api._fits(-1, **ok)
