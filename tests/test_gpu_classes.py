"""The TW_Engine-compatible C++ classes driven the way TW_Core drives them
(FFTCCTWorkerThread::processFrameFFTCC_ICGN, fftcctworkerthread.cpp:399-572):
tw_frame_loop runs cuFFTCC2D + cuICGN2D over a frame sequence with a
reference reset and the POI tracking around it (cuAccumulateUV,
cuUpdatePOIpos, cuAccumulatePOI), and every frame's results are checked
against the oracle run on the same reference, target and POIs."""
import subprocess

import numpy as np
import pytest

import oracle
from tests import cases
from tw_amd import api

pytestmark = pytest.mark.gpu

EXE = api.LIB_PATH.parent / "tw_frame_loop"


def _frames(raw, n, count):
    per = n * (12 + 4 + 1 + 24 + 8 + 8)
    assert len(raw) == count * per
    out = []
    for k in range(count):
        b = raw[k * per:(k + 1) * per]
        o = 0

        def take(nbytes, dt, shape=None):
            nonlocal o
            a = np.frombuffer(b[o:o + nbytes], dt)
            o += nbytes
            return a.reshape(shape) if shape else a
        f = dict(u=take(4 * n, np.float32), v=take(4 * n, np.float32), zncc=take(4 * n, np.float32),
                 iters=take(4 * n, np.int32), flags=take(n, np.uint8), p6=take(24 * n, np.float32, (n, 6)),
                 cur=take(8 * n, np.int32, (n, 2)), acc_u=take(4 * n, np.float32), acc_v=take(4 * n, np.float32))
        out.append(f)
    return out


def test_frame_loop_against_oracle(tmp_path):
    W = H = 384
    d1 = api.affine(1.7, -2.2, 0.003, -0.002, 0.001, 191.5, 191.5)
    d2 = api.affine(-3.1, 0.8, -0.002, 0.004, 0.0, 191.5, 191.5)
    ref, t1 = api.synth_pair(W, H, 2200, 31, d1)
    _, t2 = api.synth_pair(W, H, 2200, 31, d2, want_ref=False)
    # period 2: frame 3 (t1) becomes the reference; the POIs advance by the
    # integer part of frame 2's result; frame 4 (t2) runs at the moved POIs
    frames = [t1, t2, t1, t2]
    roi, S, G, M, N = (1, 1, W - 2, H - 2), 12, 9, 14, 32
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
    poi0 = case.poi()
    n = len(poi0)
    res = _frames(out.read_bytes(), n, len(frames))
    # what TW_Core's reset does to the POIs and the accumulators (oracle)
    moved = oracle.track_update_poi(poi0, res[1]["u"], res[1]["v"])
    acc_u, acc_v = oracle.track_accumulate_uv(np.zeros(n), np.zeros(n), res[1]["u"], res[1]["v"])
    setups = [(ref, poi0), (ref, poi0), (t1, moved), (t1, moved)]
    for k, (tar, g) in enumerate(zip(frames, res)):
        R, poi = setups[k]
        ex = oracle.fftcc_exact(R, tar, poi, N)
        f32 = oracle.fftcc(R, tar, poi, N)
        # FFT-CC ZNCC: 1e-6 of the exact value, 3e-5 of the FP32 path
        assert np.max(np.abs(g["zncc"] - ex["zncc"])) < 1e-6, k
        assert np.max(np.abs(g["zncc"] - f32["zncc"])) < 3e-5, k
        u0, v0 = ex["u"].astype(np.float32), ex["v"].astype(np.float32)
        o = oracle.icgn(R, tar, roi, S, S, poi, u0, v0)
        o64 = oracle.icgn(R, tar, roi, S, S, poi, u0, v0, precision="f64")
        assert np.array_equal(g["flags"], o["flags"]), k
        same = (g["iters"] == o["iters"]) & (g["iters"] == o64["iters"]) & (o["flags"] == 0)
        assert same.mean() > 0.99, (k, same.mean())
        gi = [1, 2, 4, 5]
        # u, v: within 5e-5 px of the FP64 oracle and 1e-4 px of the FP32 oracle
        assert np.max(np.abs(g["u"][same] - o64["u"][same])) < 5e-5, k
        assert np.max(np.abs(g["v"][same] - o64["v"][same])) < 5e-5, k
        assert np.max(np.abs(g["u"][same] - o["u"][same])) < 1e-4, k
        assert np.max(np.abs(g["v"][same] - o["v"][same])) < 1e-4, k
        # displacement gradients: on these 25 x 25 subsets the reference's own
        # FP32 path sits up to 2.1e-6 from the FP64 oracle (both stop once the
        # weighted |dp| < 1e-3, so the final iterate carries that noise); the
        # HIP path, FP32 warp arithmetic with FP64 sums, must be no further from
        # FP64 than 1.5x the reference's own distance
        own = np.abs(o["p6"][same][:, gi] - o64["p6"][same][:, gi])
        d64 = np.abs(g["p6"][same][:, gi] - o64["p6"][same][:, gi])
        assert d64.max() <= 1.5 * own.max() + 1e-7, (k, d64.max(), own.max())
        assert np.percentile(d64, 99) <= 1.5 * np.percentile(own, 99) + 1e-7, k
        d32 = np.abs(g["p6"][same][:, gi] - o["p6"][same][:, gi])
        assert d32.max() <= 2.5 * own.max() + 1e-7, k
        # POI tracking: current POIs and accumulators, bit-exact
        assert np.array_equal(g["cur"], oracle.track_current_poi(poi, g["u"], g["v"])), k
        if k < 2:
            assert not g["acc_u"].any() and not g["acc_v"].any()
        else:
            assert np.array_equal(g["acc_u"], acc_u) and np.array_equal(g["acc_v"], acc_v), k
    # the reset frame is the new reference itself: zero update at iteration 1
    assert np.all(res[2]["u"] == 0) and np.all(res[2]["v"] == 0) and np.all(res[2]["iters"] == 1)
    # the POIs did move (frame 2's u is about -3 px)
    assert np.mean(moved[:, 1] - poi0[:, 1]) < -1.5


WORKER = api.LIB_PATH.parent / "tw_icgn_worker"


def test_icgn_worker_host_arrays_against_oracle(tmp_path):
    """TW_Core's IC-GN worker (ICGNWorkerThread::processFrame,
    icgnworkerthread.cpp:45-73) with cuICGN2D in place of ICGN2D_CPU: per
    frame ResetRefImg, then ICGN2D_Algorithm on host arrays of POIs and
    integer guesses.  Two frames, the second on a new reference.  Bars as in
    the parity tests: flags equal; u, v within 1e-4 px of the FP32 oracle and
    5e-5 px of the FP64 oracle where the iteration counts agree; counts differ
    only where the oracle stopped with |dp| within 5 % of delta_p, at most
    max(2, 0.2 %) POIs or three times the FP32 oracle's own flips against
    FP64, + 2."""
    W = H = 320
    d1 = api.affine(2.4, -1.6, 0.004, -0.003, 0.001, 159.5, 159.5)
    d2 = api.affine(-1.2, 3.3, -0.002, 0.002, 0.0, 159.5, 159.5)
    ref, t1 = api.synth_pair(W, H, 1500, 41, d1)
    _, t2 = api.synth_pair(W, H, 1500, 41, d2, want_ref=False)
    roi, S, G, M, N = (1, 1, W - 2, H - 2), 15, 7, 10, 32
    case = cases.Case("worker", W, S, G, N, roi, M, ref, None, None)
    poi = case.poi()
    n = len(poi)
    nx, ny = len(np.unique(poi[:, 1])), len(np.unique(poi[:, 0]))
    assert nx * ny == n
    frames = [(ref, t1), (t1, t2)]
    guesses = []
    for R, T in frames:
        ex = oracle.fftcc_exact(R, T, poi, N)
        guesses.append((ex["u"].astype(np.float32), ex["v"].astype(np.float32)))
    (tmp_path / "poi.bin").write_bytes(np.ascontiguousarray(poi, np.int32).tobytes())
    (tmp_path / "uv0.bin").write_bytes(b"".join(u.tobytes() + v.tobytes() for u, v in guesses))
    paths = []
    for k, (R, T) in enumerate(frames):
        for tag, img in (("r", R), ("t", T)):
            p = tmp_path / f"{tag}{k}.raw"
            img.tofile(p)
            paths.append(str(p))
    out = tmp_path / "out.bin"
    r = subprocess.run([str(WORKER), str(W), str(H), *map(str, roi), str(S), str(nx), str(ny),
                        str(tmp_path / "poi.bin"), str(tmp_path / "uv0.bin"), str(out), *paths],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    per = 13 * n
    assert len(raw) == per * len(frames)
    for k, ((R, T), (u0, v0)) in enumerate(zip(frames, guesses)):
        b = raw[k * per:(k + 1) * per]
        gu = np.frombuffer(b[:4 * n], np.float32)
        gv = np.frombuffer(b[4 * n:8 * n], np.float32)
        gi = np.frombuffer(b[8 * n:12 * n], np.int32)
        gf = np.frombuffer(b[12 * n:], np.uint8)
        o = oracle.icgn(R, T, roi, S, S, poi, u0, v0)
        o64 = oracle.icgn(R, T, roi, S, S, poi, u0.astype(np.float64), v0.astype(np.float64), precision="f64")
        assert np.array_equal(gf, o["flags"]), k
        assert np.mean(gf == 0) > 0.95, k
        longer = gi > o["iters"]
        assert np.all(np.abs(o["dpnorm"][longer] / 1e-3 - 1.0) < 0.05), k
        own = int(np.count_nonzero(o["iters"] != o64["iters"]))     # the FP32 path's own flips
        assert np.count_nonzero(gi != o["iters"]) <= max(2, n // 500, 3 * own + 2), k
        ok = (gi == o["iters"]) & (o["flags"] == 0)
        assert np.max(np.abs(gu[ok] - o["u"][ok])) <= 1e-4 and np.max(np.abs(gv[ok] - o["v"][ok])) <= 1e-4, k
        ok64 = ok & (gi == o64["iters"])
        assert np.max(np.abs(gu[ok64] - o64["u"][ok64])) <= 5e-5, k
        assert np.max(np.abs(gv[ok64] - o64["v"][ok64])) <= 5e-5, k
