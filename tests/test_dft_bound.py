"""The error bound behind the fp16 matrix-core DFT FFT-CC engine (N = 64).

That engine computes the circular correlation C[a][b] = sum r[y][x]
t[(y + a) % 64][(x + b) % 64] of the mean-centred windows approximately, in
fp16 DFTs that use the real-input (Hermitian) half spectrum, and then
evaluates exactly, in integers, every shift whose approximate value is within
2 x 1.25 x kDftBound ||r|| ||t|| / 64^2 of the approximate maximum.  The
result is the exact argmax (lowest index on ties) only if

    |C^ - C| <= kDftBound ||r|| ||t|| / 64^2   at every shift.

This file recomputes that constant from first principles: the fp16 DFT table
the kernel uploads (rebuilt here the same way), its spectral-norm and
entrywise errors, the fp16 / fp32 unit roundoffs, and a standard summation
model (any fp32 sum of n terms errs by at most gamma_{n-1} sum |terms|) for
each pass, with the absolute error of fp16 values below the normal range
taken as 2^-14 (flushed to zero) in the worst case.  It checks the constant
compiled into the kernel covers that worst case, that the table identities the
half-spectrum packing relies on hold exactly, and that the operands stay in
fp16 range.  A matrix-level numpy model of the same arithmetic then checks the
bound on windows chosen to stress it (all energy in the packed DC/Nyquist
columns, exact ties, a single-pixel window with ||r|| = 1).  The model is not
the kernel (its fp32 sums run in BLAS order); the kernel itself is held to the
integer engine and the oracle bit for bit by the GPU tests.
"""
import re
from pathlib import Path

import numpy as np

N = 64
ROOT = Path(__file__).resolve().parents[1]


def fp16_table():
    """F[j][y] = exp(-2 pi i ((j y) mod 64) / 64) rounded to fp16 per
    component, as the kernel's table upload builds it."""
    e = (np.arange(N)[:, None] * np.arange(N)[None, :]) & 63
    ang = -2.0 * np.pi * e / 64.0
    re_ = np.cos(ang).astype(np.float16).astype(np.float64)
    im_ = np.sin(ang).astype(np.float16).astype(np.float64)
    return re_ + 1j * im_, np.exp(1j * ang)


def table_constants():
    Ft, F = fp16_table()
    D = Ft - F
    return (float(np.linalg.norm(D, 2)), float(np.abs(D).max()),
            float(max(np.linalg.norm(Ft, axis=0).max(), np.linalg.norm(Ft, axis=1).max())))


def dft_bound(flush):
    """N^2 max |C^ - C| / (||r|| ||t||) for the half-spectrum engine.

    Passes (their fp32 sums have 64, 128, 128 and 64 terms per component):
      A  Y = r G, G = the first 32 columns of F with column 0 = F[:,0] + i F[:,32]
         (both exact in fp16), rounded to fp16;
      B  W = F Y; column 0 unpacked into R^[:,0], R^[:,32] with the j <-> -j
         partners (fp32); P = conj(R^) T^ s rounded to fp16, column 0 packed as
         P[:,0] + i P[:,32];
      C  Q = P^T conj(F) for k = 0 .. 31, scaled by 2^-8 and by the weights
         1 (k = 0) / 2 (k >= 1), rounded to fp16;
      D  C^T = Re(conj(F) Q) over k = 0 .. 31, the imaginary table part of
         k = 0 (zero in F) set to (-1)^b, so that Q[32][a] (row 0's imaginary
         part) enters as (-1)^b Q[32][a].
    Completing the half spectrum by conjugation gives exactly the full-spectrum
    computation with one set of errors; the packed column's two halves get
    errors whose squared norms sum to that of the packed column (parallelogram
    law), so the forward bound is the full-spectrum one.  The packed column
    costs a factor 2 in the fp16 rounding of P and of Q's row 0, taken here
    for all columns (worst case: all energy in columns 0 and 32).  ||r|| >= 1
    and ||t|| >= 1 (non-zero integer arrays) turn the absolute sub-normal terms
    into relative ones; s ||r|| ||t|| 4096 >= 2^13 gives 1/s <= ||r|| ||t|| / 2.
    """
    delta, dmax, phi = table_constants()
    u, v = 2.0 ** -11, 2.0 ** -24
    g63, g127 = (n * v / (1 - n * v) for n in (63, 127))
    tiny = 2.0 ** -14 if flush else 2.0 ** -25
    r2 = np.sqrt(2.0)
    # forward, per unit ||r||: ||Y~ - r F||_F <= g, ||R^~ - F r F||_F <= h
    g = delta + 8 * phi * g63 + u * (8 + delta + 8 * phi * g63) + 64 * r2 * tiny
    h = 8 * g + delta * (8 + g) + (8 * r2 * phi * g127 + v * (8 + delta)) * (8 + g)
    dP = 2 * 64 * h + h * h                     # ||conj(R^~) T^~ - conj(R^) T^||_1, ||R^||_F = 64 ||r||
    fp32P = 6 * v * (64 + h) ** 2               # the fp32 products and the unpacking
    pi0 = 4096 + dP + fp32P                     # ||P^||_1 before the fp16 rounding
    sinv = 0.5
    roundP = 2 * u * pi0 + 4096 * r2 * tiny * sinv
    pi = pi0 + roundP                           # ||P~||_1, the inverse's input
    inv = pi * (2 * dmax * (1 + dmax)           # table error in C and in D
                + 2 * g127 * (1 + dmax) ** 2    # C's sums
                + u * (1 + dmax) ** 2 * (1 + 2 * g127)   # Q to fp16
                + g63 * (1 + dmax) ** 2 * (1 + u))       # D's sums
    inv += 64 * r2 * tiny * 256 * sinv * (1 + dmax)      # Q's sub-normal floor
    return dP + fp32P + roundP + inv


def kernel_constant():
    src = (ROOT / "tw_amd" / "csrc" / "fftcc.hip").read_text()
    m = re.search(r"constexpr double kDftBound = ([0-9.]+);", src)
    assert m, "kDftBound not found"
    return float(m.group(1))


def test_table_identities():
    Ft, F = fp16_table()
    jm = (-np.arange(N)) % N
    # F~[-j][y] = conj F~[j][y] exactly: the unpacking and the k <-> 64 - k
    # pairing of the half spectrum hold for the rounded table
    assert np.array_equal(Ft[jm, :], np.conj(Ft))
    # columns 0 and 32 are exact (1 and (-1)^y): the packed first-pass column
    assert np.array_equal(Ft[:, 0], np.ones(N))
    assert np.array_equal(Ft[:, 32], (-1.0) ** np.arange(N))
    # entries are 0 or normal fp16 (>= 0.098 in magnitude): no sub-normal operand
    mags = np.abs(np.concatenate([Ft.real.ravel(), Ft.imag.ravel()]))
    assert mags[mags > 0].min() > 2.0 ** -14


def test_constants_and_kernel_bound():
    delta, dmax, phi = table_constants()
    assert 1.2e-3 < delta < 1.35e-3 and 1.6e-4 < dmax < 1.8e-4 and 8.0 < phi < 8.001
    b_keep, b_flush = dft_bound(False), dft_bound(True)
    print(f"bound: {b_keep:.3f} (sub-normals kept), {b_flush:.3f} (flushed); kernel {kernel_constant()}")
    assert 14.0 < b_keep < b_flush < 23.0
    k = kernel_constant()
    assert b_flush <= k <= 1.1 * b_flush


def test_operand_ranges():
    # |r|, |t| <= 255 integers: first-pass components <= 64 * 255 (fp16 max 65504)
    assert 64 * 255 < 65504
    # |s P| <= s ||R^||_inf ||T^||_inf <= s (64 ||r||)(64 ||t||) < 2^14; the packed
    # column's components are sums of two such values
    assert 2 * 2 ** 14 * (1 + 0.01) < 65504
    # Q rows k >= 1 carry weight 2 / 256 over 64 terms of modulus <= 2^14 sqrt(2)
    assert 64 * 2 ** 14 * np.sqrt(2) * 2 / 256 < 65504


# ---------------------------------------------------------------------------
# matrix-level model of the same arithmetic
def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def model_half(r, t):
    Ft, _ = fp16_table()
    Fr, Fi = Ft.real.astype(np.float32), Ft.imag.astype(np.float32)
    G_r, G_i = Fr[:, :32].copy(), Fi[:, :32].copy()
    G_i[:, 0] = (-1.0) ** np.arange(N)
    nr = float(np.sqrt((r.astype(np.float64) ** 2).sum()))
    nt = float(np.sqrt((t.astype(np.float64) ** 2).sum()))
    _, ex = np.frexp(4096.0 * nr * nt)
    s = np.float32(2.0 ** (14 - int(ex)))

    def forward(x):
        x = x.astype(np.float32)
        yr, yi = _f16(x @ G_r), _f16(x @ G_i)
        wr = np.concatenate([Fr, -Fi], 1) @ np.concatenate([yr, yi], 0)
        wi = np.concatenate([Fr, Fi], 1) @ np.concatenate([yi, yr], 0)
        return wr.astype(np.float32), wi.astype(np.float32)

    Rr, Ri = forward(r)
    Tr, Ti = forward(t)
    pr = (Rr * Tr + Ri * Ti).astype(np.float32)
    pi = (Rr * Ti - Ri * Tr).astype(np.float32)
    jm = (-np.arange(N)) % N
    h = np.float32(0.5)
    ar, ai = h * (Rr[:, 0] + Rr[jm, 0]), h * (Ri[:, 0] - Ri[jm, 0])
    br, bi = h * (Ri[:, 0] + Ri[jm, 0]), h * (Rr[jm, 0] - Rr[:, 0])
    cr, ci = h * (Tr[:, 0] + Tr[jm, 0]), h * (Ti[:, 0] - Ti[jm, 0])
    dr, di = h * (Ti[:, 0] + Ti[jm, 0]), h * (Tr[jm, 0] - Tr[:, 0])
    p0r, p0i = ar * cr + ai * ci, ar * ci - ai * cr
    p1r, p1i = br * dr + bi * di, br * di - bi * dr
    pr[:, 0], pi[:, 0] = p0r - p1i, p0i + p1r
    Pr, Pi = _f16(pr * s), _f16(pi * s)
    qr = (np.concatenate([Pr.T, Pi.T], 1) @ np.concatenate([Fr, Fi], 0)).astype(np.float32)
    qi = (np.concatenate([Pi.T, Pr.T], 1) @ np.concatenate([Fr, -Fi], 0)).astype(np.float32)
    w = np.full((32, 1), 2.0 / 256, np.float32)
    w[0] = 1.0 / 256
    Qr, Qi = _f16(qr * w), _f16(qi * w)
    FiD = Fi[:, :32].copy()
    FiD[:, 0] = (-1.0) ** np.arange(N)
    ct = (np.concatenate([Fr[:, :32], FiD], 1) @ np.concatenate([Qr, Qi], 0)).astype(np.float32)
    return ct.T.astype(np.float64) * 256.0 / float(s) / 4096.0, nr * nt


def exact_corr(r, t):
    R = np.fft.fft2(r.astype(np.float64))
    T = np.fft.fft2(t.astype(np.float64))
    return np.real(np.fft.ifft2(np.conj(R) * T))


def _windows():
    rng = np.random.default_rng(64)
    y, x = np.mgrid[0:N, 0:N]
    out = {}
    noise = rng.integers(0, 256, (N, N))
    k = np.ones(5) / 5
    blur = np.apply_along_axis(lambda v: np.convolve(np.tile(v, 3), k, "same")[N:2 * N], 1, noise)
    blur = np.apply_along_axis(lambda v: np.convolve(np.tile(v, 3), k, "same")[N:2 * N], 0, blur)
    out["noise"] = (noise, np.roll(noise, (3, -7), (0, 1)))
    out["speckle"] = (np.rint(blur * 2 - 128).clip(0, 255), np.rint(np.roll(blur, (5, 2), (0, 1)) * 2 - 128).clip(0, 255))
    rows = rng.integers(0, 256, N)
    out["dc column"] = (np.tile(rows[:, None], (1, N)), np.tile(np.roll(rows, 9)[:, None], (1, N)))
    alt = (-1) ** x
    out["nyquist column"] = (128 + (rows[:, None] // 2) * alt, 128 + (np.roll(rows, 4)[:, None] // 2) * alt)
    out["dc + nyquist"] = (np.clip(rows[:, None] + 60 * alt, 0, 255), np.clip(np.roll(rows, 1)[:, None] - 60 * alt, 0, 255))
    out["checkerboard"] = (128 + 127 * (-1) ** (x + y), 128 + 127 * (-1) ** (x + y + 1))
    one = np.full((N, N), 100)
    one2 = one.copy()
    one2[17, 40] = 101
    out["one pixel"] = (one2, np.roll(one2, (2, 2), (0, 1)))
    return out


def test_model_within_bound_and_candidates_hold_the_argmax():
    bound = dft_bound(False)
    for name, (R, T) in _windows().items():
        R, T = np.asarray(R, np.int64), np.asarray(T, np.int64)
        a = (int(R.sum()) + 2048) >> 12
        b = (int(T.sum()) + 2048) >> 12
        r, t = R - a, T - b
        C, rho = model_half(r, t)
        E = exact_corr(r, t)
        err = float(np.abs(C - E).max())
        lim = bound * rho / 4096
        print(f"{name}: max |C^ - C| = {err:.3g}, bound {lim:.3g} ({err / lim:.3f} of it)")
        assert err <= lim, name
        # every exact maximiser is among the candidates the kernel evaluates
        cand = C.ravel() >= C.max() - 2 * 1.25 * lim
        best = np.isclose(E.ravel(), E.max(), rtol=0, atol=1e-6)
        assert np.all(cand[best]), name
