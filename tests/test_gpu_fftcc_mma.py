"""Edge cases of the matrix-core FFT-CC engine against the exact-integer
oracle (brute-force correlation, oracle/tw_oracle.c:orc_fftcc_exact): exact
ties (lowest row-major index wins, TW_paDIC_FFTCC2D_CPU.cpp:300-311),
saturated pixels (the signed bytes R - 128 and T - 128 at both extremes), every
window edge class (one 32 x 32 tile, padded tiles, the full 64 x 64 layout),
flat windows and windows that leave the image.  Bars: u, v and status equal;
ZNCC within 1e-6 of the exact value."""
import numpy as np
import pytest

import oracle
from tests import cases
from tests.gpu_helpers import ctx_for, run_fftcc
from tw_amd import api

pytestmark = pytest.mark.gpu

SIZES = [2, 4, 6, 8, 20, 30, 32, 34, 40, 62, 64]


def _case(ref, tar):
    h, w = ref.shape
    return cases.Case("edge", w, 8, 8, 32, (1, 1, w - 2, h - 2), 2, ref, tar, None)


def _check(ref, tar, poi, N):
    with ctx_for(_case(ref, tar), N=N) as ctx:
        g = run_fftcc(ctx, poi)
    ex = oracle.fftcc_exact(ref, tar, poi, N)
    assert np.array_equal(g["status"], ex["status"]), N
    ok = ex["status"] == oracle.FFTCC_OK
    assert np.array_equal(g["u"][ok].astype(np.int32), ex["u"][ok]), N
    assert np.array_equal(g["v"][ok].astype(np.int32), ex["v"][ok]), N
    if ok.any():
        assert np.max(np.abs(g["zncc"][ok] - ex["zncc"][ok])) <= 1e-6, N
    return g, ex


@pytest.mark.parametrize("N", SIZES)
def test_exact_ties_lowest_index(N):
    """Each window holds one bright reference pixel and two (or three) equal
    bright target pixels, so the integer correlation ties exactly at two or
    three shifts; the lowest row-major index must win."""
    rng = np.random.default_rng(N)
    size = 4 * 64 + 8
    ref = np.full((size, size), 10, np.uint8)
    tar = np.full((size, size), 10, np.uint8)
    poi = []
    for iy in range(3):
        for ix in range(3):
            py, px = 40 + 80 * iy, 40 + 80 * ix
            y0, x0 = py - N // 2, px - N // 2
            ref[y0, x0] = 200
            for _ in range(2 + (iy + ix) % 2):
                ky, kx = rng.integers(0, N, 2)
                tar[y0 + ky, x0 + kx] = 200
            poi.append((py, px))
    poi = np.array(poi, np.int32)
    g, ex = _check(ref, tar, poi, N)
    assert np.all(ex["status"] == oracle.FFTCC_OK)


@pytest.mark.parametrize("N", SIZES)
def test_saturated_and_random_windows(N):
    """Binary 0/255 images reach both ends of the signed-byte range; a
    speckle-free random image adds many near-equal shifts."""
    rng = np.random.default_rng(100 + N)
    for ref, tar in (
        ((rng.random((160, 160)) < 0.5) * 255, (rng.random((160, 160)) < 0.5) * 255),
        (rng.integers(0, 256, (160, 160)), rng.integers(0, 256, (160, 160))),
    ):
        ref = ref.astype(np.uint8)
        tar = np.roll(ref, (2, -3), (0, 1)) if N >= 8 else tar.astype(np.uint8)
        iy, ix = np.meshgrid(np.arange(40, 121, 20), np.arange(40, 121, 20), indexing="ij")
        poi = np.stack([iy.ravel(), ix.ravel()], 1).astype(np.int32)
        _check(ref, tar, poi, N)


@pytest.mark.parametrize("N", [4, 30, 64])
def test_flat_and_outside_windows(N):
    """A constant window is FLAT, a window that leaves the image is OUT; both
    leave u, v and ZNCC untouched, as the CPU path does.  (The oracle reads
    whole windows, so it is asked only about the windows inside the image.)"""
    rng = np.random.default_rng(5)
    ref = rng.integers(0, 256, (200, 200)).astype(np.uint8)
    tar = np.roll(ref, (1, 1), (0, 1))
    ref[60:130, 60:130] = 77                     # flat reference window at (95, 95)
    poi = np.array([[95, 95], [150, 40], [N // 2, N // 2], [N // 2 - 1, 100], [100, 200 - N // 2 + 1]], np.int32)
    with ctx_for(_case(ref, tar), N=N) as ctx:
        g = run_fftcc(ctx, poi)
    assert list(g["status"]) == [api.FFTCC_FLAT, api.FFTCC_OK, api.FFTCC_OK, api.FFTCC_OUT, api.FFTCC_OUT]
    for i in (0, 3, 4):
        assert g["u"][i] == 0 and g["v"][i] == 0 and g["zncc"][i] == 0
    _check(ref, tar, poi[:3], N)


@pytest.mark.parametrize("N", [2, 6, 30, 34, 40, 46, 48, 50, 62])
def test_every_row_alignment_and_image_corner(N):
    """Windows whose rows start at each of the 16 byte offsets within an
    aligned 16-byte chunk (the engine reads window rows as the aligned chunks
    that hold them; the odd image width also varies the offset from row to
    row), and one window that ends at the image's last pixel."""
    rng = np.random.default_rng(300 + N)
    h, w = 2 * N + 40, 2 * N + 57
    ref = rng.integers(0, 256, (h, w)).astype(np.uint8)
    tar = np.roll(ref, (1, -2), (0, 1))
    poi = [(N // 2 + 3 + i % 3, N // 2 + i) for i in range(16)] + [(h - N // 2, w - N // 2)]
    g, ex = _check(ref, tar, np.array(poi, np.int32), N)
    assert np.all(ex["status"] == oracle.FFTCC_OK)
