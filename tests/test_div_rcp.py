"""The warp update's quotients (icgn.hip div_rcp): a / b from y = RN(1/b) by
Markstein's correction, q0 = RN(a y), r = RN(a - b q0) (one FMA, exact),
q = RN(q0 + r y) (one FMA), and q0 itself for a zero numerator.  The FP32
operations are emulated exactly on the CPU (each product, remainder and
corrected sum is exact in float64 before its single rounding to float32)
and the result must equal IEEE float32 division bit for bit, signed zeros
included, over the warp's ranges and wider."""
import numpy as np

f32 = np.float32


def div_rcp(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    y = (f32(1) / b).astype(np.float32)
    q0 = (a * y).astype(np.float32)
    r = (a.astype(np.float64) - b.astype(np.float64) * q0.astype(np.float64)).astype(np.float32)
    q = (q0.astype(np.float64) + r.astype(np.float64) * y.astype(np.float64)).astype(np.float32)
    return np.where(a == 0, q0, q)


def check(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    got = div_rcp(a, b).view(np.uint32)
    ref = (a / b).astype(np.float32).view(np.uint32)
    bad = got != ref
    assert not bad.any(), (a[bad][:4], b[bad][:4])


def test_determinant_like_denominators():
    rng = np.random.default_rng(3)
    for sign in (1, -1):
        b = sign * (1 + rng.normal(0, 0.05, 2_000_000))
        a = rng.uniform(-1, 1, 2_000_000) * 2.0 ** rng.integers(-40, 40, 2_000_000)
        check(a, b)


def test_wide_denominators_and_exact_quotients():
    rng = np.random.default_rng(4)
    b = rng.uniform(0.5, 2, 2_000_000) * 2.0 ** rng.integers(-20, 20, 2_000_000)
    a = rng.uniform(-1, 1, 2_000_000) * 2.0 ** rng.integers(-40, 40, 2_000_000)
    check(a, b)
    b = (1 + rng.normal(0, 0.05, 1_000_000)).astype(np.float32)
    k = rng.integers(-2 ** 24, 2 ** 24, 1_000_000).astype(np.float32)
    check((k.astype(np.float64) * b.astype(np.float64)).astype(np.float32), b)


def test_signed_zero_numerators():
    check([0.0, -0.0, 0.0, -0.0], [1.1, 1.1, -0.9, -0.9])
