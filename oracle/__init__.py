"""CPU oracle for the paDIC hot path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/libtworacle.so (a plain-C restatement of TW_Engine's
CPU path, see tw_oracle.h).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this package, and only as the checker
or as the timed CPU baseline.  The product (tw_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libtworacle.so"

SUCCESS, DARK_SUBSET, SINGULAR_HESSIAN, SINGULAR_WARP, OUT_OF_ROI = range(5)
FFTCC_OK, FFTCC_FLAT = 0, 1

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "libtworacle.so"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp, i, f, d = C.c_void_p, C.c_int, C.c_float, C.c_double
        sig = {
            "orc_num_poi": ([i] * 8 + [C.POINTER(C.c_int), C.POINTER(C.c_int)], i),
            "orc_poi_grid": ([i] * 10 + [vp], None),
            "orc_gradients_f32": ([vp, i, i, i, i, i, i, i, vp, vp, vp], i),
            "orc_bicubic_lut_f32": ([vp, i, i, i, vp, vp, vp, i, i, i, i, vp], i),
            "orc_fftcc_f32": ([vp, vp, i, vp, i, i, vp, vp, vp, vp, vp, i], i),
            "orc_fftcc_f64": ([vp, vp, i, vp, i, i, vp, vp, vp, vp, vp, i], i),
            "orc_fftcc_exact": ([vp, vp, i, vp, i, i, vp, vp, vp, vp, i], i),
            "orc_icgn_f32": ([vp, vp, i, i, i, i, i, i, i, i, i, i, f, vp, i, vp, vp, vp, vp, vp, vp, vp, i], i),
            "orc_icgn_f64": ([vp, vp, i, i, i, i, i, i, i, i, i, i, d, vp, i, vp, vp, vp, vp, vp, vp, vp, i], i),
            "orc_icgn_interp_f32": ([vp, vp, i, i, i, i, i, i, i, i, i, i, f, vp, i, vp, vp, vp, vp, vp, vp, vp, i, i],
                                    i),
            "orc_icgn_interp_f64": ([vp, vp, i, i, i, i, i, i, i, i, i, i, d, vp, i, vp, vp, vp, vp, vp, vp, vp, i, i],
                                    i),
            "orc_bspline_lut_f32": ([vp, i, i, i, i, i, i, i, vp], i),
            "orc_fft1d_f32": ([vp, i, i], None),
            "orc_fft1d_f64": ([vp, i, i], None),
            "orc_potrf6_f32": ([vp], i),
            "orc_potrs6_f32": ([vp, vp], None),
            "orc_potrf6_f64": ([vp], i),
            "orc_potrs6_f64": ([vp, vp], None),
            "orc_lut_eval_f32": ([vp, f, f], f),
            "orc_warp_update_f32": ([vp, vp], i),
            "orc_warp_update_f64": ([vp, vp], i),
            "orc_track_update_poi": ([vp, i, vp, vp], None),
            "orc_track_current_poi": ([vp, i, vp, vp, vp], None),
            "orc_track_accumulate_uv": ([i, vp, vp, vp, vp], None),
            "orc_track_current_poi_acc": ([vp, i, vp, vp, vp, vp, vp], None),
            "orc_strain_ls": ([vp, i, i, i, i, vp, vp], None),
        }
        for name, (a, r) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def num_poi(roi_w, roi_h, sx, sy, gx, gy, mx, my):
    nx, ny = C.c_int(), C.c_int()
    ok = lib().orc_num_poi(roi_w, roi_h, sx, sy, gx, gy, mx, my, C.byref(nx), C.byref(ny))
    return nx.value, ny.value, bool(ok)


def poi_grid(start_x, start_y, nx, ny, mx, my, sx, sy, gx, gy) -> np.ndarray:
    out = np.zeros((ny * nx, 2), np.int32)
    lib().orc_poi_grid(start_x, start_y, nx, ny, mx, my, sx, sy, gx, gy, _p(out))
    return out


def gradients(img, start_x, start_y, roi_w, roi_h, with_xy=True):
    img = np.ascontiguousarray(img, np.uint8)
    gx = np.zeros((roi_h, roi_w), np.float32)
    gy = np.zeros_like(gx)
    gxy = np.zeros_like(gx) if with_xy else None
    rc = lib().orc_gradients_f32(_p(img), img.shape[1], img.shape[1], img.shape[0], start_x, start_y,
                                 roi_w, roi_h, _p(gx), _p(gy), _p(gxy))
    if rc:
        raise ValueError("gradient frame outside the image")
    return gx, gy, gxy


def bicubic_lut(img, start_x, start_y, roi_w, roi_h) -> np.ndarray:
    img = np.ascontiguousarray(img, np.uint8)
    tx, ty, txy = gradients(img, start_x, start_y, roi_w, roi_h)
    lut = np.zeros((roi_h, roi_w, 16), np.float32)
    rc = lib().orc_bicubic_lut_f32(_p(img), img.shape[1], img.shape[1], img.shape[0], _p(tx), _p(ty), _p(txy),
                                   start_x, start_y, roi_w, roi_h, _p(lut))
    if rc:
        raise ValueError("LUT bounds")
    return lut


def fftcc(ref, tar, poi_yx, N, precision="f32", nthreads=0):
    ref = np.ascontiguousarray(ref, np.uint8)
    tar = np.ascontiguousarray(tar, np.uint8)
    poi = np.ascontiguousarray(poi_yx, np.int32)
    n = len(poi)
    dt = np.float32 if precision == "f32" else np.float64
    u = np.full(n, np.nan, dt)
    v = np.full(n, np.nan, dt)
    z = np.full(n, np.nan, dt)
    gap = np.full(n, np.nan, dt)
    st = np.full(n, 255, np.uint8)
    fn = lib().orc_fftcc_f32 if precision == "f32" else lib().orc_fftcc_f64
    rc = fn(_p(ref), _p(tar), ref.shape[1], _p(poi), n, N, _p(u), _p(v), _p(z), _p(st), _p(gap), nthreads)
    if rc:
        raise ValueError("bad N")
    return dict(u=u, v=v, zncc=z, status=st, gap=gap)


def fftcc_exact(ref, tar, poi_yx, N, nthreads=0):
    ref = np.ascontiguousarray(ref, np.uint8)
    tar = np.ascontiguousarray(tar, np.uint8)
    poi = np.ascontiguousarray(poi_yx, np.int32)
    n = len(poi)
    u = np.zeros(n, np.int32)
    v = np.zeros(n, np.int32)
    z = np.full(n, np.nan, np.float64)
    st = np.full(n, 255, np.uint8)
    rc = lib().orc_fftcc_exact(_p(ref), _p(tar), ref.shape[1], _p(poi), n, N, _p(u), _p(v), _p(z), _p(st), nthreads)
    if rc:
        raise ValueError("bad N")
    return dict(u=u, v=v, zncc=z, status=st)


def icgn(ref, tar, roi, sx, sy, poi_yx, u0, v0, max_iter=20, delta_p=0.001, precision="f32", nthreads=0,
         iters_init=-1, interp="bicubic"):
    """roi = (start_x, start_y, roi_w, roi_h); interp "bicubic" or "bspline".
    Returns dict of per-POI outputs."""
    ref = np.ascontiguousarray(ref, np.uint8)
    tar = np.ascontiguousarray(tar, np.uint8)
    poi = np.ascontiguousarray(poi_yx, np.int32)
    n = len(poi)
    dt = np.float32 if precision == "f32" else np.float64
    u = np.array(u0, dt, copy=True).reshape(n)
    v = np.array(v0, dt, copy=True).reshape(n)
    p6 = np.zeros((n, 6), dt)
    iters = np.full(n, iters_init, np.int32)
    flags = np.full(n, 255, np.uint8)
    dpn = np.zeros(n, np.float64)
    times = np.zeros(2, np.float64)
    fn = lib().orc_icgn_interp_f32 if precision == "f32" else lib().orc_icgn_interp_f64
    h, w = ref.shape
    rc = fn(_p(ref), _p(tar), w, w, h, roi[0], roi[1], roi[2], roi[3], sx, sy, max_iter, delta_p, _p(poi), n,
            _p(u), _p(v), _p(p6), _p(iters), _p(flags), _p(dpn), _p(times), nthreads,
            {"bicubic": 0, "bspline": 1}[interp])
    if rc:
        raise ValueError("bad geometry")
    return dict(u=u, v=v, p6=p6, iters=iters, flags=flags, dpnorm=dpn, times_ms=times)


def fft1d(x: np.ndarray, sign: int = -1, precision="f64") -> np.ndarray:
    dt = np.complex64 if precision == "f32" else np.complex128
    a = np.ascontiguousarray(x, dt).copy()
    (lib().orc_fft1d_f32 if precision == "f32" else lib().orc_fft1d_f64)(_p(a), len(a), sign)
    return a


def cholesky_solve(H: np.ndarray, b: np.ndarray, precision="f64"):
    dt = np.float32 if precision == "f32" else np.float64
    a = np.ascontiguousarray(H, dt).copy()
    rhs = np.ascontiguousarray(b, dt).copy()
    potrf = lib().orc_potrf6_f32 if precision == "f32" else lib().orc_potrf6_f64
    potrs = lib().orc_potrs6_f32 if precision == "f32" else lib().orc_potrs6_f64
    info = potrf(_p(a))
    potrs(_p(a), _p(rhs))
    return info, np.tril(a), rhs


def warp_update(p, dp, precision="f64"):
    dt = np.float32 if precision == "f32" else np.float64
    a = np.ascontiguousarray(p, dt).copy()
    d = np.ascontiguousarray(dp, dt)
    fn = lib().orc_warp_update_f32 if precision == "f32" else lib().orc_warp_update_f64
    singular = fn(_p(a), _p(d))
    return a, bool(singular)


def lut_eval(cell16: np.ndarray, tx: float, ty: float) -> float:
    c = np.ascontiguousarray(cell16, np.float32)
    return float(lib().orc_lut_eval_f32(_p(c), tx, ty))


def track_update_poi(poi_yx, u, v) -> np.ndarray:
    """Copy of poi_yx moved by (int(v), int(u)) (TW_Core updatePOIpos_Kernel)."""
    out = np.array(poi_yx, np.int32, copy=True, order="C")
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    lib().orc_track_update_poi(_p(out), len(out), _p(u), _p(v))
    return out


def track_current_poi(poi_yx, u, v) -> np.ndarray:
    poi = np.ascontiguousarray(poi_yx, np.int32)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.empty_like(poi)
    lib().orc_track_current_poi(_p(poi), len(poi), _p(u), _p(v), _p(out))
    return out


def track_current_poi_acc(poi_yx, u, v, acc_u, acc_v) -> np.ndarray:
    """The 7-argument cuAccumulatePOI (accumulatePOINew_Kernel)."""
    poi = np.ascontiguousarray(poi_yx, np.int32)
    u, v, au, av = (np.ascontiguousarray(a, np.float32) for a in (u, v, acc_u, acc_v))
    out = np.empty_like(poi)
    lib().orc_track_current_poi_acc(_p(poi), len(poi), _p(u), _p(v), _p(au), _p(av), _p(out))
    return out


def track_accumulate_uv(acc_u, acc_v, u, v):
    """Returns (acc_u + u, acc_v + v) computed as the reference does, in FP32."""
    au = np.array(acc_u, np.float32, copy=True)
    av = np.array(acc_v, np.float32, copy=True)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    lib().orc_track_accumulate_uv(len(au), _p(u), _p(v), _p(au), _p(av))
    return au, av


def strain_ls(field, grid, M):
    """(d/dx, d/dy) of field[ny][nx] by the least-squares window of
    calcustrain4_x_0_L1_5_40.m, per pixel."""
    f = np.ascontiguousarray(field, np.float32)
    ny, nx = f.shape
    gx = np.empty((ny - 2 * M, nx - 2 * M), np.float64)
    gy = np.empty_like(gx)
    lib().orc_strain_ls(_p(f), nx, ny, grid, M, _p(gx), _p(gy))
    return gx, gy


def bspline_lut(img, start_x, start_y, roi_w, roi_h) -> np.ndarray:
    """BicubicSplineCoefficients_s as a [roi_h][roi_w][16] LUT (4k+l ~ y^k x^l)."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    lut = np.zeros((roi_h, roi_w, 16), np.float32)
    if lib().orc_bspline_lut_f32(_p(img), w, w, h, start_x, start_y, roi_w, roi_h, _p(lut)):
        raise ValueError("the ROI needs a 1-pixel frame left/top and 2 pixels right/bottom")
    return lut
