"""The fp16 matrix-core DFT FFT-CC engine (N = 64, TWDIC_FFTCC_VARIANT=9):
an approximate correlation surface from fp16 2-D DFTs, then the exact
integer correlation at every shift within the surface's proven error of its
maximum.  Its peaks must be the exact integer argmax with the CPU path's
lowest-index tie-break (TW_paDIC_FFTCC2D_CPU.cpp:300-311), so every output
(u, v, ZNCC, status) must equal the integer matrix-core engine's bit for bit,
and both must equal the exact-integer oracle: on the 2048^2 workload's
157,609 POIs, on exact ties, saturated and random windows, near-uniform
windows (more candidate shifts than the engine re-evaluates, so it falls
back to the integer engine for them), flat windows and windows that leave
the image."""
import numpy as np
import pytest

import oracle
from tests import cases
from tests.gpu_helpers import ctx_for, run_fftcc
from tw_amd import api

pytestmark = pytest.mark.gpu

OUTS = ("u", "v", "zncc", "status")


@pytest.fixture
def dft(monkeypatch):
    monkeypatch.setenv("TWDIC_FFTCC_VARIANT", "9")


def _both(case, poi, monkeypatch, N=64):
    monkeypatch.setenv("TWDIC_FFTCC_VARIANT", "9")
    with ctx_for(case, N=N) as ctx:
        g9 = run_fftcc(ctx, poi)
    monkeypatch.setenv("TWDIC_FFTCC_VARIANT", "8")
    with ctx_for(case, N=N) as ctx:
        g8 = run_fftcc(ctx, poi)
    return g9, g8


def _same(a, b):
    for k in OUTS:
        assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), k


def _case(ref, tar):
    h, w = ref.shape
    return cases.Case("edge", w, 8, 8, 64, (1, 1, w - 2, h - 2), 2, ref, tar, None)


def _exact(ref, tar, poi, g):
    ex = oracle.fftcc_exact(ref, tar, poi, 64)
    assert np.array_equal(g["status"], ex["status"])
    ok = ex["status"] == oracle.FFTCC_OK
    assert np.array_equal(g["u"][ok].astype(np.int32), ex["u"][ok])
    assert np.array_equal(g["v"][ok].astype(np.int32), ex["v"][ok])
    if ok.any():
        assert np.max(np.abs(g["zncc"][ok] - ex["zncc"][ok])) <= 1e-6


def test_full_2048_frame_equals_integer_engine(monkeypatch):
    c = cases.speckle_2048()
    poi = c.poi()
    print(f"[gpu-test] dft engine, 2048^2 frame: {len(poi)} POIs", flush=True)
    g9, g8 = _both(c, poi, monkeypatch)
    _same(g9, g8)
    assert np.all(g9["status"] == api.FFTCC_OK)
    rng = np.random.default_rng(64)
    idx = np.unique(np.concatenate([rng.choice(len(poi), 1500, replace=False), [0, len(poi) - 1]]))
    _exact(c.ref, c.tar, np.ascontiguousarray(poi[idx]), {k: g9[k][idx] for k in OUTS})


def test_ties_saturated_random_nearly_uniform(monkeypatch):
    rng = np.random.default_rng(9)
    size = 4 * 64 + 8
    # exact ties: one bright reference pixel, two or three equal bright target pixels per window
    ref = np.full((size, size), 10, np.uint8)
    tar = np.full((size, size), 10, np.uint8)
    poi = []
    for iy in range(3):
        for ix in range(3):
            py, px = 40 + 80 * iy, 40 + 80 * ix
            y0, x0 = py - 32, px - 32
            ref[y0, x0] = 200
            for _ in range(2 + (iy + ix) % 2):
                ky, kx = rng.integers(0, 64, 2)
                tar[y0 + ky, x0 + kx] = 200
            poi.append((py, px))
    sets = [(ref, tar, np.array(poi, np.int32))]
    iy, ix = np.meshgrid(np.arange(40, 121, 20), np.arange(40, 121, 20), indexing="ij")
    grid = np.stack([iy.ravel(), ix.ravel()], 1).astype(np.int32)
    b = ((rng.random((160, 160)) < 0.5) * 255).astype(np.uint8)
    sets.append((b, np.roll(b, (2, -3), (0, 1)), grid))                      # saturated 0/255
    r = rng.integers(0, 256, (160, 160)).astype(np.uint8)
    sets.append((r, rng.integers(0, 256, (160, 160)).astype(np.uint8), grid))  # unrelated random windows
    # near-uniform, not flat: a constant field with a few +-1 pixels, so the
    # whole surface lies within the error bound (the integer-engine fallback)
    u = np.full((160, 160), 120, np.uint8)
    for _ in range(40):
        y, x = rng.integers(0, 160, 2)
        u[y, x] += rng.choice([-1, 1])
    sets.append((u, np.roll(u, (5, 7), (0, 1)), grid))
    for k, (R, T, P) in enumerate(sets):
        print(f"[gpu-test] dft engine, edge set {k}", flush=True)
        g9, g8 = _both(_case(R, T), P, monkeypatch)
        _same(g9, g8)
        _exact(R, T, P, g9)


def test_flat_and_outside_windows(monkeypatch):
    rng = np.random.default_rng(5)
    ref = rng.integers(0, 256, (200, 200)).astype(np.uint8)
    tar = np.roll(ref, (1, 1), (0, 1))
    ref[60:130, 60:130] = 77
    poi = np.array([[95, 95], [150, 40], [32, 32], [31, 100], [100, 169]], np.int32)
    g9, g8 = _both(_case(ref, tar), poi, monkeypatch)
    _same(g9, g8)
    assert list(g9["status"]) == [api.FFTCC_FLAT, api.FFTCC_OK, api.FFTCC_OK, api.FFTCC_OUT, api.FFTCC_OUT]
