"""The TW_Engine-compatible C++ classes driven the way TW_Core drives them
(FFTCCTWorkerThread::processFrameFFTCC_ICGN, fftcctworkerthread.cpp:399-572):
tw_frame_loop runs cuFFTCC2D + cuICGN2D over a frame sequence with a
reference reset, and every frame's results are checked against the oracle."""
import subprocess

import numpy as np
import pytest

import oracle
from tests import cases
from tw_amd import api

pytestmark = pytest.mark.gpu

EXE = api.LIB_PATH.parent / "tw_frame_loop"


def test_frame_loop_against_oracle(tmp_path):
    W = H = 384
    d1 = api.affine(1.7, -2.2, 0.003, -0.002, 0.001, 191.5, 191.5)
    d2 = api.affine(-3.1, 0.8, -0.002, 0.004, 0.0, 191.5, 191.5)
    ref, t1 = api.synth_pair(W, H, 2200, 31, d1)
    _, t2 = api.synth_pair(W, H, 2200, 31, d2, want_ref=False)
    frames = [t1, t2, t1]                     # frame 3 becomes the new reference (period 2)
    roi, S, G, M, N = (1, 1, W - 2, H - 2), 12, 9, 10, 32
    paths = []
    for i, img in enumerate([ref] + frames):
        p = tmp_path / f"f{i}.raw"
        img.tofile(p)
        paths.append(str(p))
    out = tmp_path / "out.bin"
    r = subprocess.run([str(EXE), str(W), str(H), *map(str, roi), str(S), str(G), str(M), str(N), "2", str(out),
                        *paths], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    case = cases.Case("loop", W, S, G, N, roi, M, ref, None, None)
    poi = case.poi()
    n = len(poi)
    raw = out.read_bytes()
    per = n * (4 * 4 + 1 + 24)
    assert len(raw) == 3 * per
    refs = [ref, ref, t1]
    for k, tar in enumerate(frames):
        b = raw[k * per:(k + 1) * per]
        u, v, z = (np.frombuffer(b[i * 4 * n:(i + 1) * 4 * n], np.float32) for i in range(3))
        it = np.frombuffer(b[12 * n:16 * n], np.int32)
        fl = np.frombuffer(b[16 * n:17 * n], np.uint8)
        p6 = np.frombuffer(b[17 * n:], np.float32).reshape(n, 6)
        R = refs[k]
        ex = oracle.fftcc_exact(R, tar, poi, N)
        f32 = oracle.fftcc(R, tar, poi, N)
        assert np.max(np.abs(z - ex["zncc"])) < 1e-6
        assert np.max(np.abs(z - f32["zncc"])) < 3e-5
        o = oracle.icgn(R, tar, roi, S, S, poi, ex["u"].astype(np.float32), ex["v"].astype(np.float32))
        assert np.array_equal(fl, o["flags"])
        same = (it == o["iters"]) & (o["flags"] == 0)
        assert same.mean() > 0.99
        assert np.max(np.abs(u[same] - o["u"][same])) < 1e-4
        assert np.max(np.abs(v[same] - o["v"][same])) < 1e-4
        assert np.max(np.abs(p6[same][:, [1, 2, 4, 5]] - o["p6"][same][:, [1, 2, 4, 5]])) < 3e-6
        if k == 2:                            # reference reset to this very frame
            assert np.all(u == 0) and np.all(v == 0) and np.all(it == 1)
