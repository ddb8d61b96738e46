"""Bicubic B-spline interpolation (ICGN2DInterpolationFLag::BicubicSpline):
the oracle's restatement of BicubicSplineCoefficients_s (TW_utils.cpp:
521-602) pinned to a numpy matrix formulation and to the identities the
interpolant must keep; the HIP LUT bit-exact to it and IC-GN in B-spline mode
held to the parity bars."""
import numpy as np
import pytest

import oracle
from tests import cases
from tw_amd import api

CP = np.array([[71, -19, 5, -1], [-19, 95, -25, 5], [5, -25, 95, -19], [-1, 5, -19, 71]]) / 56.0
B = np.array([[-1, 3, -3, 1], [3, -6, 3, 0], [-3, 0, 3, 0], [1, 4, 1, 0]]) / 6.0


def _numpy_lut(img, i, j):
    om = img[i:i + 4, j:j + 4].astype(np.float64)        # rows y-1..y+2 of ROI pixel (i, j) with a 1-px frame
    beta = CP @ om.T @ CP.T
    return (B @ beta.T @ B.T)[::-1, ::-1].ravel()


def _eval(cell, tx, ty):
    return sum(cell[4 * k + l] * ty ** k * tx ** l for k in range(4) for l in range(4))


def test_oracle_bspline_lut_pins():
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (48, 56)).astype(np.uint8)
    lut = oracle.bspline_lut(img, 1, 1, 52, 44)
    for (i, j) in [(0, 0), (5, 7), (43, 51), (20, 30)]:
        want = _numpy_lut(img, i, j)
        assert np.max(np.abs(lut[i, j] - want)) <= 2e-5 * max(1.0, np.abs(want).max()), (i, j)
    # a constant image is reproduced (the prefilter rows sum to 1, B-spline partition of unity)
    c = np.full((20, 20), 117, np.uint8)
    lc = oracle.bspline_lut(c, 1, 1, 16, 16)
    for tx, ty in [(0.0, 0.0), (0.3, 0.7), (0.99, 0.01)]:
        assert abs(_eval(lc[4, 5].astype(np.float64), tx, ty) - 117.0) < 1e-3
    # the interpolant passes through the pixels: (CP rows 0+4*1+2)/6 and (1+4*2+3)/6
    # are unit vectors, so each cell's corners give the four neighbouring pixels;
    # a transposed layout (y^l x^k) or a missing 180-degree turn fails here
    for (i, j) in [(0, 0), (7, 3), (42, 50)]:
        cell = lut[i, j].astype(np.float64)
        y, x = 1 + i, 1 + j
        for tx, ty in [(0, 0), (1, 0), (0, 1), (1, 1)]:
            assert abs(_eval(cell, tx, ty) - float(img[y + ty, x + tx])) < 2e-3, (i, j, tx, ty)
    # halfway between two pixels of a linear ramp the symmetric prefilter is exact
    yy, xx = np.mgrid[0:20, 0:20]
    ramp = (3 * xx + 5 * yy + 10).astype(np.uint8)
    lr = oracle.bspline_lut(ramp, 1, 1, 16, 16)
    for tx, ty in [(0.5, 0.0), (0.0, 0.5), (0.5, 0.5)]:
        x, y = 1 + 6 + tx, 1 + 4 + ty
        assert abs(_eval(lr[4, 6].astype(np.float64), tx, ty) - (3 * x + 5 * y + 10)) < 1e-3
    # the frame rule: 2 pixels right/below the ROI
    with pytest.raises(ValueError):
        oracle.bspline_lut(img, 1, 1, 54, 44)


def test_oracle_icgn_bspline_identities():
    c = cases.speckle_512()
    roi = (1, 1, 508, 508)
    poi = c.poi()[:300]
    z = np.zeros(len(poi), np.float32)
    o = oracle.icgn(c.ref, c.ref, roi, 15, 15, poi, z, z, interp="bspline")
    assert np.all(o["flags"] == 0) and np.all(o["iters"] == 1) and not o["u"].any() and not o["v"].any()
    ex = oracle.fftcc_exact(c.ref, c.tar, poi, 32)
    o = oracle.icgn(c.ref, c.tar, roi, 15, 15, poi, ex["u"].astype(np.float32), ex["v"].astype(np.float32),
                    interp="bspline")
    ok = o["flags"] == 0
    assert ok.mean() > 0.99
    # the resampled target moved by (2.35, -1.72): B-spline IC-GN recovers it
    assert np.median(np.abs(o["u"][ok] - 2.35)) < 0.02 and np.median(np.abs(o["v"][ok] + 1.72)) < 0.02


@pytest.mark.gpu
def test_gpu_bspline_lut_bit_exact_and_icgn_parity():
    from tests.gpu_helpers import check_icgn, run_icgn
    c = cases.speckle_512()
    roi = (1, 1, 508, 508)
    cfg = api.make_config(512, 512, *roi, 15, 15, 10, 10, 3, 3, 32, interp=api.INTERP_BSPLINE)
    with api.Context(cfg) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        ctx.precompute_target()
        ctx.sync()
        lut = ctx.download_buffer(api.BUF_LUT, (508, 508, 16), np.float32)
        assert np.array_equal(lut, oracle.bspline_lut(c.tar, *roi))
        poi = ctx.poi_grid().download()
        ex = oracle.fftcc_exact(c.ref, c.tar, poi, 32)
        u0, v0 = ex["u"].astype(np.float32), ex["v"].astype(np.float32)
        g = run_icgn(ctx, poi, u0, v0)
    o = oracle.icgn(c.ref, c.tar, roi, 15, 15, poi, u0, v0, interp="bspline")
    o64 = oracle.icgn(c.ref, c.tar, roi, 15, 15, poi, u0.astype(np.float64), v0.astype(np.float64),
                      precision="f64", interp="bspline")
    # same bars as the bicubic LUT mode: u, v within 1e-4 px of the FP32 path and
    # 5e-5 of FP64; gradients within 3e-6 / 2e-6.  Iteration counts: every flip
    # must sit in the |dp| noise band; against FP64 their number is held to
    # three times the FP32 CPU path's own flips against FP64, plus 2 (the B-spline
    # interpolant converges more slowly, so more POIs end near delta_p)
    check_icgn(g, o, gtol=3e-6)
    own = int(np.sum(o["iters"] != o64["iters"]))
    print(f"  FP32 oracle vs FP64 oracle: {own} iteration-count flips")
    check_icgn(g, o64, utol=5e-5, gtol=2e-6, max_flips=3 * own + 2)
