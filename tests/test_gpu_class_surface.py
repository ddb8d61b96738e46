"""cuFFTCC2D's host-output API and its ROI-only constructor, driven through
the compiled class layer (tw_fftcc_host) as a TW_Core caller would.

* InitializeFFTCC / ComputeFFTCC / DestroyFFTCC (TW_paDIC_cuFFTCC2D.cu:
  237-461): the callee allocates hcreateptr arrays [NumPOIY][NumPOIX]; the
  results read through their row pointers must equal the exact-integer
  oracle (peaks bit-exact, ZNCC within 1e-6) on every POI of two frames, and
  DestroyFFTCC must release and null the arrays.
* Constructors: the whole-image one (ROI size before ROI start,
  TW_paDIC_cuFFTCC2D.cu:192-210) places the POIs in image coordinates; the
  ROI-only one (:178-190) treats the images as the ROI and places them
  ROI-relative (:254-266; TW_cuUtils.cu:83-100).  Both grids must equal the
  oracle's, and the two constructors must agree on the same pixels."""
import subprocess

import numpy as np
import pytest

import oracle
from tw_amd import api

pytestmark = pytest.mark.gpu

EXE = api.LIB_PATH.parent / "tw_fftcc_host"
W, H = 400, 360
ROI = (17, 23, 300, 280)          # start x, start y, width, height
S, G, M = 12, 9, 3                # FFT-CC window 2S = 24, as the reference


def _run(args, tmp_path, ny_nx_frames):
    out = tmp_path / "out.bin"
    r = subprocess.run([str(EXE), *map(str, args[:-1]), str(out), *args[-1]], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint8)
    nx, ny = np.frombuffer(raw[:8].tobytes(), np.int32)
    n = int(nx) * int(ny)
    poi = np.frombuffer(raw[8:8 + 8 * n].tobytes(), np.int32).reshape(n, 2)
    rest = np.frombuffer(raw[8 + 8 * n:].tobytes(), np.float32)
    assert rest.size == ny_nx_frames * 3 * n
    fr = rest.reshape(ny_nx_frames, 3, n)
    return int(nx), int(ny), poi, fr


def test_host_output_api_both_constructors(tmp_path):
    d1 = api.affine(2.6, -1.4, 0.004, -0.002, 0.001, W / 2, H / 2)
    d2 = api.affine(-3.2, 2.1, -0.003, 0.002, 0.0, W / 2, H / 2)
    ref, t1 = api.synth_pair(W, H, 1800, 41, d1)
    _, t2 = api.synth_pair(W, H, 1800, 41, d2, want_ref=False)
    x0, y0, rw, rh = ROI
    paths = {}
    for name, img in (("ref", ref), ("t1", t1), ("t2", t2)):
        full = tmp_path / f"{name}.raw"
        img.tofile(full)
        crop = tmp_path / f"{name}_roi.raw"
        np.ascontiguousarray(img[y0:y0 + rh, x0:x0 + rw]).tofile(crop)
        paths[name] = (str(full), str(crop))
    nx_o, ny_o, ok = oracle.num_poi(rw, rh, S, S, G, G, M, M)
    assert ok
    # whole-image constructor
    nx, ny, poi, fr = _run(["whole", W, H, x0, y0, rw, rh, S, G, M,
                            [paths["ref"][0], paths["t1"][0], paths["t2"][0]]], tmp_path, 2)
    assert (nx, ny) == (nx_o, ny_o)
    assert np.array_equal(poi, oracle.poi_grid(x0, y0, nx, ny, M, M, S, S, G, G))
    for k, tar in enumerate((t1, t2)):
        ex = oracle.fftcc_exact(ref, tar, poi, 2 * S)
        assert np.all(ex["status"] == oracle.FFTCC_OK)
        assert np.array_equal(fr[k, 0], ex["u"]) and np.array_equal(fr[k, 1], ex["v"]), k
        assert np.max(np.abs(fr[k, 2] - ex["zncc"])) < 1e-6, k
    # ROI-only constructor: the images are the ROI crops, POIs ROI-relative
    nx2, ny2, poi2, fr2 = _run(["roi", rw, rh, S, G, M, [paths["ref"][1], paths["t1"][1], paths["t2"][1]]],
                               tmp_path, 2)
    assert (nx2, ny2) == (nx_o, ny_o)
    assert np.array_equal(poi2, oracle.poi_grid(0, 0, nx2, ny2, M, M, S, S, G, G))
    assert np.array_equal(poi2 + np.array([y0, x0], np.int32), poi)
    # the same pixels: both constructors give the same peaks and ZNCC
    assert np.array_equal(fr2.view(np.uint32), fr.view(np.uint32))
