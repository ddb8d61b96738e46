"""The fused IC-GN interpolation's row-polynomial form (CPU checks).

The target's bicubic LUT (BicubicCoefficients_s, TW_utils.cpp:604-700, built
by the oracle from central-difference corner derivatives) is, per cell, the
tensor-product Catmull-Rom spline of the 4x4 neighbourhood.  The fused mode
evaluates that spline as
    R_k(tx) = c0 + tx/2 (c1 + tx (c2 + tx c3))          for rows k = 0..3,
    c0 = p1, c1 = p2 - p0, c2 = 2 p0 - 5 p1 + 4 p2 - p3, c3 = 3 (p1 - p2) + p3 - p0,
    T = R1 + ty/2 (d1 + ty (d2 + ty d3)),
    d1 = R2 - R0, d2 = 2 R0 - 5 R1 + 4 R2 - R3, d3 = 3 (R1 - R2) + R3 - R0,
with the row coefficients stored as fp16 c0, c1/2, c2/2, c3/2 (precompute.hip
k_row_poly; icgn.hip interp_row_poly evaluates c0 + tx (c1/2 + tx (c2/2 +
tx c3/2)), which rounds exactly as the form above: every intermediate is half
of it).  These tests check, on the CPU:
  * the form equals the oracle's LUT polynomial exactly (exact dyadic
    arithmetic in float64, so a wrong coefficient or sign fails);
  * the stored coefficients are exact in fp16 (|c| <= 1530, halves of
    multiples of 1/2 below 766);
  * the FP32 evaluation in the kernel's operation order (every FMA rounded
    once, emulated exactly with fractions) returns a uniform neighbourhood's
    value exactly, which keeps the reference's DarkSubset behaviour on flat
    target regions, and stays within 2e-4 of the exact value elsewhere."""
from fractions import Fraction

import numpy as np

import oracle


def row_coeffs(p):
    p0, p1, p2, p3 = (int(v) for v in p)
    return p1, p2 - p0, 2 * p0 - 5 * p1 + 4 * p2 - p3, 3 * (p1 - p2) + p3 - p0


def row_poly_exact(nb, tx, ty):
    R = []
    for k in range(4):
        c0, c1, c2, c3 = row_coeffs(nb[k])
        R.append(c0 + tx / 2 * (c1 + tx * (c2 + tx * c3)))
    d1 = R[2] - R[0]
    d2 = 2 * R[0] - 5 * R[1] + 4 * R[2] - R[3]
    d3 = 3 * (R[1] - R[2]) + R[3] - R[0]
    return R[1] + ty / 2 * (d1 + ty * (d2 + ty * d3))


def f32(x: float) -> Fraction:
    return Fraction(float(np.float32(x)))


def _round32(q: Fraction) -> Fraction:
    """Round an exact rational to the nearest float32 (ties to even)."""
    if q == 0:
        return Fraction(0)
    v = float(q)                          # nearest double
    lo = np.float32(v)
    # step to the float32 neighbours and pick the nearest by exact comparison
    cands = {lo, np.nextafter(lo, np.float32(np.inf)), np.nextafter(lo, np.float32(-np.inf))}
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - q), int(np.float32(c).view(np.uint32)) & 1))
    return Fraction(float(best))


def fma32(a, b, c):
    return _round32(Fraction(a) * Fraction(b) + Fraction(c))


def mul32(a, b):
    return _round32(Fraction(a) * Fraction(b))


def sub32(a, b):
    return _round32(Fraction(a) - Fraction(b))


def row_poly_f32(nb, tx, ty):
    """The kernel's FP32 operation order (interp_row_poly)."""
    tx, ty = f32(tx), f32(ty)
    hx = mul32(Fraction(1, 2), tx)
    R = []
    for k in range(4):
        c0, c1, c2, c3 = row_coeffs(nb[k])
        h1, h2, h3 = Fraction(c1, 2), Fraction(c2, 2), Fraction(c3, 2)
        r = fma32(tx, fma32(tx, fma32(tx, h3, h2), h1), c0)
        # the same value as the unhalved order with t/2 applied last
        assert r == fma32(hx, fma32(fma32(tx, c3, c2), tx, c1), c0)
        R.append(r)
    d1 = sub32(R[2], R[0])
    d2 = fma32(4, R[2], fma32(-5, R[1], fma32(2, R[0], -R[3])))
    d3 = fma32(3, sub32(R[1], R[2]), sub32(R[3], R[0]))
    return fma32(mul32(Fraction(1, 2), ty), fma32(fma32(d3, ty, d2), ty, d1), R[1])


def lut_poly_exact(cell, tx, ty):
    a = [Fraction(float(v)) for v in cell]
    return sum(a[4 * k + l] * tx ** l * ty ** k for k in range(4) for l in range(4))


def test_row_polynomials_equal_the_bicubic_lut_exactly():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (24, 24), dtype=np.uint8)
    sx0, sy0, rw, rh = 2, 2, 18, 18
    lut = oracle.bicubic_lut(img, sx0, sy0, rw, rh)
    for _ in range(300):
        i, j = rng.integers(0, rh - 1), rng.integers(0, rw - 1)       # interior cells
        y, x = sy0 + i, sx0 + j
        nb = img[y - 1:y + 3, x - 1:x + 3]
        tx = Fraction(int(rng.integers(0, 1024)), 1024)
        ty = Fraction(int(rng.integers(0, 1024)), 1024)
        assert row_poly_exact(nb, tx, ty) == lut_poly_exact(lut[i, j], tx, ty)


def test_row_coefficients_are_exact_in_fp16():
    # extremes over all byte rows: |c2| <= 1530, |c3| <= 1020, all integers
    ext = [np.array(v) for v in ([0, 255, 0, 255], [255, 0, 255, 0], [255, 0, 0, 255], [0, 255, 255, 0],
                                 [255, 255, 0, 0], [0, 0, 255, 255])]
    rng = np.random.default_rng(1)
    rows = ext + [rng.integers(0, 256, 4) for _ in range(2000)]
    for r in rows:
        c = row_coeffs(r)
        assert all(abs(v) <= 1530 for v in c)
        for v in (c[0], c[1] / 2, c[2] / 2, c[3] / 2):
            assert float(np.float16(v)) == v


def test_fp32_order_is_exact_on_uniform_neighbourhoods_and_close_elsewhere():
    rng = np.random.default_rng(9)
    for v in (0, 1, 17, 128, 254, 255):
        nb = np.full((4, 4), v, np.uint8)
        for _ in range(20):
            tx, ty = float(np.float32(rng.random())), float(np.float32(rng.random()))
            assert row_poly_f32(nb, tx, ty) == v
    worst = Fraction(0)
    for _ in range(200):
        nb = rng.integers(0, 256, (4, 4))
        tx, ty = float(np.float32(rng.random())), float(np.float32(rng.random()))
        exact = row_poly_exact(nb, Fraction(tx), Fraction(ty))
        worst = max(worst, abs(row_poly_f32(nb, tx, ty) - exact))
    assert worst < Fraction(2, 10000), float(worst)
