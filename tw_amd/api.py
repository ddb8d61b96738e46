"""ctypes binding of the C ABI in include/twdic.h (libtwdic.so, built in-tree).

This is plumbing for tests and benchmarks: every call goes straight to the
HIP library.  There is no CPU fallback; if the library is missing or no GPU is
usable, the calls raise.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libtwdic.so"

# status codes (twdic_status)
OK, EINVAL, ENOMEM, EBOUNDS, EHIP, EBUSY, ENODEV = range(7)
# ICGN2DFlag
ICGN_SUCCESS, ICGN_DARK_SUBSET, ICGN_SINGULAR_HESSIAN, ICGN_SINGULAR_WARP, ICGN_OUT_OF_ROI = range(5)
FFTCC_OK, FFTCC_FLAT, FFTCC_OUT = range(3)
INTERP_LUT, INTERP_FUSED, INTERP_BSPLINE = 0, 1, 2
BUF_RX, BUF_RY, BUF_TX, BUF_TY, BUF_TXY, BUF_LUT, BUF_REF, BUF_TAR = range(8)
DEFORM_AFFINE, DEFORM_SINUSOID, DEFORM_RESAMPLE = range(3)


class TwdicError(RuntimeError):
    pass


class Config(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("img_w", "img_h", "roi_x", "roi_y", "roi_w", "roi_h", "sx", "sy",
                                       "gx", "gy", "mx", "my", "fft_n", "max_iter")] + [
        ("delta_p", C.c_float), ("interp", C.c_int)]


class Deform(C.Structure):
    _fields_ = [("kind", C.c_int)] + [(n, C.c_double) for n in (
        "u0", "v0", "exx", "eyy", "exy", "xc", "yc", "amp_u", "per_u", "amp_v", "per_v")]


class Speckle(C.Structure):
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("nspots", C.c_int), ("seed", C.c_uint64),
                ("sigma_min", C.c_double), ("sigma_max", C.c_double), ("amp_min", C.c_double),
                ("amp_max", C.c_double), ("noise_sigma", C.c_double), ("noise_seed", C.c_uint64),
                ("boxes", C.c_void_p), ("nboxes", C.c_int)]


class Pair(C.Structure):
    _fields_ = [("d_ref", C.c_void_p), ("d_tar", C.c_void_p), ("stride", C.c_int), ("img_w", C.c_int),
                ("img_h", C.c_int), ("roi_x", C.c_int), ("roi_y", C.c_int), ("roi_w", C.c_int), ("roi_h", C.c_int),
                ("sx", C.c_int), ("sy", C.c_int), ("fft_n", C.c_int), ("max_iter", C.c_int), ("delta_p", C.c_float),
                ("d_poi_yx", C.c_void_p), ("n", C.c_int), ("d_u", C.c_void_p), ("d_v", C.c_void_p),
                ("d_zncc", C.c_void_p), ("d_p6", C.c_void_p), ("d_iters", C.c_void_p), ("d_flags", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("icgn_chunk", C.c_int), ("fftcc_variant", C.c_int), ("use_graph", C.c_int),
                ("icgn_waves", C.c_int), ("icgn_xcd_ranges", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("slots", C.c_int), ("queue_capacity", C.c_int), ("hbm_budget", C.c_size_t),
                ("max_width", C.c_int), ("max_height", C.c_int), ("max_pois", C.c_int)]


class Request(C.Structure):
    _fields_ = [("ref_id", C.c_int), ("poiset", C.c_int), ("target", C.c_void_p), ("stride", C.c_int),
                ("subset", C.c_int), ("fft_n", C.c_int), ("max_iter", C.c_int), ("delta_p", C.c_float),
                ("u", C.c_void_p), ("v", C.c_void_p), ("zncc", C.c_void_p), ("p6", C.c_void_p),
                ("iters", C.c_void_p), ("flags", C.c_void_p), ("memory", C.c_int)]


MEM_PAGEABLE, MEM_PINNED, MEM_DEVICE = 0, 1, 2


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("dropped", C.c_uint64), ("failed", C.c_uint64), ("pois", C.c_double), ("lat_p50_ms", C.c_double),
                ("lat_p99_ms", C.c_double), ("lat_max_ms", C.c_double), ("wall_s", C.c_double),
                ("hbm_budget", C.c_size_t), ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t),
                ("slots_busy_peak", C.c_int), ("state_builds", C.c_uint64), ("state_hits", C.c_uint64)]


class BatchInfo(C.Structure):
    _fields_ = [("npairs", C.c_int), ("total_pois", C.c_int64), ("fftcc_items", C.c_int64 * 3),
                ("device_bytes", C.c_size_t), ("tuning", Tuning)]


_lib = None

_VP, _I, _SZ = C.c_void_p, C.c_int, C.c_size_t
_SIGS = {
    "twdic_create": ([C.POINTER(Config), _I, C.POINTER(_VP)], _I),
    "twdic_destroy": ([_VP], _I),
    "twdic_last_error": ([_VP], C.c_char_p),
    "twdic_sync": ([_VP], _I),
    "twdic_stream": ([_VP], _VP),
    "twdic_num_poi": ([C.POINTER(Config), C.POINTER(C.c_int32), C.POINTER(C.c_int32)], _I),
    "twdic_set_reference": ([_VP, _VP, _I, _I], _I),
    "twdic_set_target": ([_VP, _VP, _I, _I], _I),
    "twdic_poi_grid": ([_VP, _VP], _I),
    "twdic_fftcc": ([_VP, _VP, _I, _VP, _VP, _VP, _VP], _I),
    "twdic_icgn": ([_VP, _VP, _I, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "twdic_icgn_noalloc": ([_VP, _VP, _I, _VP, _VP, _VP, _VP, _VP], _I),
    "twdic_frame": ([_VP, _VP, _I, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "twdic_frame_capture": ([_VP, _VP, _I, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "twdic_frame_replay": ([_VP], _I),
    "twdic_reset_poi_state": ([_VP], _I),
    "twdic_get_buffer": ([_VP, _I, C.POINTER(_VP), C.POINTER(_SZ)], _I),
    "twdic_precompute_target": ([_VP], _I),
    "twdic_device_span": ([_VP, C.POINTER(_SZ)], _I),
    "twdic_malloc": ([_VP, C.POINTER(_VP), _SZ], _I),
    "twdic_free": ([_VP, _VP], _I),
    "twdic_memcpy": ([_VP, _VP, _VP, _SZ, _I], _I),
    "twdic_memset": ([_VP, _VP, _I, _SZ], _I),
    "twdic_host_alloc": ([C.POINTER(_VP), _SZ], _I),
    "twdic_host_free": ([_VP], _I),
    "twdic_track_update_poi": ([_VP, _VP, _I, _VP, _VP], _I),
    "twdic_track_current_poi": ([_VP, _VP, _I, _VP, _VP, _VP], _I),
    "twdic_track_current_poi_acc": ([_VP, _VP, _I, _VP, _VP, _VP, _VP, _VP], _I),
    "twdic_track_accumulate_uv": ([_VP, _I, _VP, _VP, _VP, _VP], _I),
    "twdic_write_csv": ([C.c_char_p, _I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP], _I),
    "twdic_strain_ls": ([_VP, _VP, _I, _I, _I, _I, _VP, _VP], _I),
    "twdic_event_record": ([_VP, _I], _I),
    "twdic_event_elapsed": ([_VP, _I, _I, C.POINTER(C.c_float)], _I),
    "twdic_batch_create": ([_I, C.POINTER(_VP)], _I),
    "twdic_batch_destroy": ([_VP], _I),
    "twdic_batch_last_error": ([_VP], C.c_char_p),
    "twdic_batch_stream": ([_VP], _VP),
    "twdic_batch_set_pairs": ([_VP, C.POINTER(Pair), _I], _I),
    "twdic_batch_set_targets": ([_VP, C.POINTER(_VP)], _I),
    "twdic_batch_set_references": ([_VP, C.POINTER(_VP)], _I),
    "twdic_batch_run": ([_VP], _I),
    "twdic_batch_run_icgn": ([_VP], _I),
    "twdic_batch_sync": ([_VP], _I),
    "twdic_batch_get_tuning": ([_VP, C.POINTER(Tuning)], _I),
    "twdic_batch_set_tuning": ([_VP, C.POINTER(Tuning)], _I),
    "twdic_batch_autotune": ([_VP, C.c_char_p, _I], _I),
    "twdic_batch_load_tuning": ([_VP, C.c_char_p, C.POINTER(_I)], _I),
    "twdic_batch_info_get": ([_VP, C.POINTER(BatchInfo)], _I),
    "twdic_batch_record": ([_VP, _I], _I),
    "twdic_batch_elapsed": ([_VP, _I, _I, C.POINTER(C.c_float)], _I),
    "twdic_server_create": ([C.POINTER(ServerConfig), C.POINTER(_VP)], _I),
    "twdic_server_destroy": ([_VP], _I),
    "twdic_server_last_error": ([_VP], C.c_char_p),
    "twdic_server_set_reference": ([_VP, _I, _VP, _I, _I, _I], _I),
    "twdic_server_add_poiset": ([_VP, _VP, _I, C.POINTER(_I)], _I),
    "twdic_server_submit": ([_VP, C.POINTER(Request), _I, C.POINTER(C.c_uint64)], _I),
    "twdic_server_poll": ([_VP, C.c_uint64, C.POINTER(_I)], _I),
    "twdic_server_wait": ([_VP, C.c_uint64], _I),
    "twdic_server_drain": ([_VP], _I),
    "twdic_server_stats_get": ([_VP, C.POINTER(ServerStats)], _I),
    "twdic_server_stats_reset": ([_VP], _I),
    "twdic_synth_pair": ([C.POINTER(Speckle), C.POINTER(Deform), _VP, _VP, _I], _I),
    "twdic_synth_displacement": ([C.POINTER(Deform), C.c_double, C.c_double, C.POINTER(C.c_double)], None),
}


def lib() -> C.CDLL:
    """Load libtwdic.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise TwdicError(f"{LIB_PATH} is missing: run __graft_entry__.build() or python tw_amd/build.py")
        L = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def device_span(ptr: int) -> int:
    """Bytes from a device pointer to the end of the HIP allocation holding it
    (twdic_device_span); raises when the runtime knows no allocation there."""
    n = C.c_size_t()
    st = lib().twdic_device_span(C.c_void_p(ptr), C.byref(n))
    if st != OK:
        raise TwdicError(f"twdic_device_span: status {st}")
    return n.value


# --------------------------------------------------------------------------- synth
def synth_pair(w: int, h: int, nspots: int, seed: int, deform: Deform | None = None, *,
               sigma=(2.0, 4.0), amp=(60.0, 200.0), noise_sigma=2.0, noise_seed: int | None = None,
               boxes: np.ndarray | None = None, want_ref=True, want_tar=True):
    """Deterministic speckle pair (reference, deformed target) as uint8 arrays."""
    sp = Speckle(w, h, nspots, seed, sigma[0], sigma[1], amp[0], amp[1], noise_sigma,
                 (seed ^ 0xA5A5A5A5) if noise_seed is None else noise_seed, None, 0)
    bx = None
    if boxes is not None and len(boxes):
        bx = np.ascontiguousarray(boxes, dtype=np.int32)
        sp.boxes = _ptr(bx)
        sp.nboxes = len(bx)
    ref = np.zeros((h, w), np.uint8) if want_ref else None
    tar = np.zeros((h, w), np.uint8) if want_tar else None
    d = deform if deform is not None else Deform(DEFORM_AFFINE, 0, 0, 0, 0, 0, w / 2, h / 2, 0, 1, 0, 1)
    rc = lib().twdic_synth_pair(C.byref(sp), C.byref(d), _ptr(ref) if ref is not None else None,
                                _ptr(tar) if tar is not None else None, w)
    if rc != 0:
        raise TwdicError("twdic_synth_pair failed")
    return ref, tar


def affine(u0, v0, exx=0.0, eyy=0.0, exy=0.0, xc=0.0, yc=0.0) -> Deform:
    return Deform(DEFORM_AFFINE, u0, v0, exx, eyy, exy, xc, yc, 0.0, 1.0, 0.0, 1.0)


def sinusoid(amp_u, per_u, amp_v, per_v) -> Deform:
    return Deform(DEFORM_SINUSOID, 0, 0, 0, 0, 0, 0, 0, amp_u, per_u, amp_v, per_v)


def resample(u0, v0) -> Deform:
    return Deform(DEFORM_RESAMPLE, u0, v0, 0, 0, 0, 0, 0, 0, 1, 0, 1)


def displacement(d: Deform, x: float, y: float) -> np.ndarray:
    out = (C.c_double * 6)()
    lib().twdic_synth_displacement(C.byref(d), x, y, out)
    return np.array(out[:])


# --------------------------------------------------------------------------- device
# Bytes per POI of each per-POI buffer the kernels index: (y, x) int32 pairs,
# float32 u / v / ZNCC / |dp| / accumulators, 6 float32 warp parameters, int32
# iteration counts, uint8 flags and FFT-CC status.
_PER_POI = {"poi": 8, "out": 8, "u": 4, "v": 4, "zncc": 4, "dpnorm": 4, "acc_u": 4, "acc_v": 4, "p6": 24,
            "iters": 4, "flags": 1, "status": 1}


def _fits(n: int, **bufs) -> None:
    """Host-side bounds check before a launch.  The C ABI takes plain pointers
    and cannot see sizes, so every per-POI buffer given here must hold n
    entries, or the call raises before anything reaches the GPU."""
    if n < 0:
        raise ValueError(f"negative POI count {n}")
    for name, a in bufs.items():
        if a is not None and a.nbytes < n * _PER_POI[name]:
            raise ValueError(f"{name}: {a.nbytes} bytes hold fewer than {n} POIs")


class DeviceArray:
    """A device allocation made through twdic_malloc, freed with the context."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx._check(lib().twdic_malloc(ctx.h, C.byref(p), max(1, self.nbytes)), "twdic_malloc")
        self.ptr = p.value
        ctx._allocs.append(self)

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        self.ctx._check(lib().twdic_memcpy(self.ctx.h, self.ptr, _ptr(a), self.nbytes, 0), "memcpy h2d")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(lib().twdic_memcpy(self.ctx.h, _ptr(out), self.ptr, self.nbytes, 1), "memcpy d2h")
        return out

    def fill(self, byte: int) -> "DeviceArray":
        self.ctx._check(lib().twdic_memset(self.ctx.h, self.ptr, byte, self.nbytes), "memset")
        return self


class Context:
    """One twdic_ctx (one HIP stream, one device)."""

    def __init__(self, cfg: Config, device: int = 0):
        self.cfg = cfg
        h = C.c_void_p()
        st = lib().twdic_create(C.byref(cfg), device, C.byref(h))
        if st != OK:
            raise TwdicError(f"twdic_create failed with status {st}")
        self.h = h.value
        self._allocs: list[DeviceArray] = []
        nx, ny = C.c_int32(), C.c_int32()
        lib().twdic_num_poi(C.byref(cfg), C.byref(nx), C.byref(ny))
        self.nx, self.ny = nx.value, ny.value

    def _check(self, st, what=""):
        if st != OK:
            msg = lib().twdic_last_error(self.h).decode() if self.h else ""
            raise TwdicError(f"{what}: status {st}: {msg}")

    def close(self):
        if self.h:
            for a in self._allocs:
                lib().twdic_free(self.h, a.ptr)
            self._allocs.clear()
            lib().twdic_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def sync(self):
        self._check(lib().twdic_sync(self.h), "sync")

    def set_reference(self, img: np.ndarray):
        img = np.ascontiguousarray(img, np.uint8)
        self._check(lib().twdic_set_reference(self.h, _ptr(img), img.shape[1], 0), "set_reference")
        self.sync()

    def set_target(self, img: np.ndarray, sync: bool = True):
        """Host image in: copied into pinned staging before the call returns
        (img may be reused at once), then one DMA on the context's stream.
        sync=False leaves that DMA in flight behind the work already queued."""
        img = np.ascontiguousarray(img, np.uint8)
        self._check(lib().twdic_set_target(self.h, _ptr(img), img.shape[1], 0), "set_target")
        if sync:
            self.sync()

    def set_reference_device(self, d: DeviceArray, stride: int):
        self._check(lib().twdic_set_reference(self.h, d.ptr, stride, 1), "set_reference")

    def set_target_device(self, d: DeviceArray, stride: int):
        self._check(lib().twdic_set_target(self.h, d.ptr, stride, 1), "set_target")

    def poi_grid(self) -> DeviceArray:
        d = self.array((self.nx * self.ny, 2), np.int32)
        self._check(lib().twdic_poi_grid(self.h, d.ptr), "poi_grid")
        return d

    def fftcc(self, poi: DeviceArray, n: int, u: DeviceArray, v: DeviceArray, zncc: DeviceArray,
              status: DeviceArray | None = None):
        _fits(n, poi=poi, u=u, v=v, zncc=zncc, status=status)
        self._check(lib().twdic_fftcc(self.h, poi.ptr, n, u.ptr, v.ptr, zncc.ptr,
                                      status.ptr if status is not None else None), "fftcc")

    def icgn(self, poi: DeviceArray, n: int, u: DeviceArray, v: DeviceArray, p6=None, iters=None,
             flags=None, dpnorm=None):
        _fits(n, poi=poi, u=u, v=v, p6=p6, iters=iters, flags=flags, dpnorm=dpnorm)
        g = lambda a: a.ptr if a is not None else None  # noqa: E731
        self._check(lib().twdic_icgn(self.h, poi.ptr, n, u.ptr, v.ptr, g(p6), g(iters), g(flags), g(dpnorm)),
                    "icgn")

    def icgn_noalloc(self, poi: DeviceArray, n: int, u: DeviceArray, v: DeviceArray, p6=None, iters=None,
                     flags=None):
        """twdic_icgn_noalloc: IC-GN with the per-POI state in the library's
        fixed per-GPU scratch (no device allocation; at most 65536 POIs)."""
        _fits(n, poi=poi, u=u, v=v, p6=p6, iters=iters, flags=flags)
        g = lambda a: a.ptr if a is not None else None  # noqa: E731
        self._check(lib().twdic_icgn_noalloc(self.h, poi.ptr, n, u.ptr, v.ptr, g(p6), g(iters), g(flags)),
                    "icgn_noalloc")

    def frame(self, poi, n, u, v, zncc, p6=None, iters=None, flags=None):
        _fits(n, poi=poi, u=u, v=v, zncc=zncc, p6=p6, iters=iters, flags=flags)
        g = lambda a: a.ptr if a is not None else None  # noqa: E731
        self._check(lib().twdic_frame(self.h, poi.ptr, n, u.ptr, v.ptr, zncc.ptr, g(p6), g(iters), g(flags)),
                    "frame")

    def frame_capture(self, poi, n, u, v, zncc, p6=None, iters=None, flags=None):
        """Record the frame step over these buffers into the context's HIP graph."""
        _fits(n, poi=poi, u=u, v=v, zncc=zncc, p6=p6, iters=iters, flags=flags)
        g = lambda a: a.ptr if a is not None else None  # noqa: E731
        self._check(lib().twdic_frame_capture(self.h, poi.ptr, n, u.ptr, v.ptr, zncc.ptr, g(p6), g(iters),
                                              g(flags)), "frame_capture")

    def frame_replay(self):
        """Launch the captured frame step on the current target."""
        self._check(lib().twdic_frame_replay(self.h), "frame_replay")

    def precompute_target(self):
        self._check(lib().twdic_precompute_target(self.h), "precompute_target")

    def buffer(self, which: int) -> tuple[int, int]:
        p, b = C.c_void_p(), C.c_size_t()
        self._check(lib().twdic_get_buffer(self.h, which, C.byref(p), C.byref(b)), "get_buffer")
        return p.value, b.value

    def download_buffer(self, which: int, shape, dtype) -> np.ndarray:
        p, b = self.buffer(which)
        out = np.empty(shape, dtype)
        assert out.nbytes <= b
        self._check(lib().twdic_memcpy(self.h, _ptr(out), p, out.nbytes, 1), "memcpy d2h")
        return out

    def track_update_poi(self, poi: DeviceArray, n: int, u: DeviceArray, v: DeviceArray):
        _fits(n, poi=poi, u=u, v=v)
        self._check(lib().twdic_track_update_poi(self.h, poi.ptr, n, u.ptr, v.ptr), "track_update_poi")

    def track_current_poi(self, poi: DeviceArray, n: int, u: DeviceArray, v: DeviceArray, out: DeviceArray):
        _fits(n, poi=poi, u=u, v=v, out=out)
        self._check(lib().twdic_track_current_poi(self.h, poi.ptr, n, u.ptr, v.ptr, out.ptr), "track_current_poi")

    def track_current_poi_acc(self, poi: DeviceArray, n: int, u: DeviceArray, v: DeviceArray, acc_u: DeviceArray,
                              acc_v: DeviceArray, out: DeviceArray):
        _fits(n, poi=poi, u=u, v=v, acc_u=acc_u, acc_v=acc_v, out=out)
        self._check(lib().twdic_track_current_poi_acc(self.h, poi.ptr, n, u.ptr, v.ptr, acc_u.ptr, acc_v.ptr, out.ptr),
                    "track_current_poi_acc")

    def track_accumulate_uv(self, n: int, u: DeviceArray, v: DeviceArray, acc_u: DeviceArray, acc_v: DeviceArray):
        _fits(n, u=u, v=v, acc_u=acc_u, acc_v=acc_v)
        self._check(lib().twdic_track_accumulate_uv(self.h, n, u.ptr, v.ptr, acc_u.ptr, acc_v.ptr),
                    "track_accumulate_uv")

    def strain_ls(self, field: DeviceArray, nx: int, ny: int, grid: int, M: int, dfdx: DeviceArray | None,
                  dfdy: DeviceArray | None):
        g = lambda a: a.ptr if a is not None else None  # noqa: E731
        self._check(lib().twdic_strain_ls(self.h, field.ptr, nx, ny, grid, M, g(dfdx), g(dfdy)), "strain_ls")

    def record(self, slot: int):
        self._check(lib().twdic_event_record(self.h, slot), "event_record")

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        self._check(lib().twdic_event_elapsed(self.h, a, b, C.byref(ms)), "event_elapsed")
        return ms.value


def make_config(img_w, img_h, roi_x, roi_y, roi_w, roi_h, sx, sy, gx, gy, mx, my, fft_n, max_iter=20,
                delta_p=0.001, interp=INTERP_FUSED) -> Config:
    return Config(img_w, img_h, roi_x, roi_y, roi_w, roi_h, sx, sy, gx, gy, mx, my, fft_n, max_iter,
                  delta_p, interp)


def num_poi(cfg: Config) -> tuple[int, int]:
    nx, ny = C.c_int32(), C.c_int32()
    lib().twdic_num_poi(C.byref(cfg), C.byref(nx), C.byref(ny))
    return nx.value, ny.value


def write_csv(path, poi_yx, int_u, int_v, p6, iters, zncc=None, fftcc_status=None, flags=None):
    """Per-POI result table in the reference's 14-column CSV layout."""
    a = lambda x, dt: None if x is None else np.ascontiguousarray(x, dt)  # noqa: E731
    poi, iu, iv = a(poi_yx, np.int32), a(int_u, np.float32), a(int_v, np.float32)
    p6, it, z = a(p6, np.float32), a(iters, np.int32), a(zncc, np.float32)
    st, fl = a(fftcc_status, np.uint8), a(flags, np.uint8)
    n = len(poi)
    g = lambda x: None if x is None else _ptr(x)  # noqa: E731
    r = lib().twdic_write_csv(str(path).encode(), n, g(poi), g(iu), g(iv), g(p6), g(it), g(z), g(st), g(fl))
    if r != OK:
        raise TwdicError(f"twdic_write_csv failed with status {r}")


class Batch:
    """A twdic_batch: many image pairs behind one handle (the direct form runs the
    single-context kernels per pair; the persistent form is off in this build).  Device
    buffers are allocated through a helper Context-free allocator owned by the
    batch object (the caller's arrays are borrowed)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().twdic_batch_create(device, C.byref(h))
        if st != OK:
            raise TwdicError(f"twdic_batch_create failed with status {st}")
        self.h = h.value
        self._keep = []

    def _check(self, st, what=""):
        if st != OK:
            raise TwdicError(f"{what}: status {st}: {lib().twdic_batch_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            lib().twdic_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_pairs(self, pairs: list):
        arr = (Pair * len(pairs))(*pairs)
        self._keep = [arr]
        self._check(lib().twdic_batch_set_pairs(self.h, arr, len(pairs)), "set_pairs")

    def set_targets(self, ptrs: list):
        arr = (C.c_void_p * len(ptrs))(*ptrs)
        self._check(lib().twdic_batch_set_targets(self.h, arr), "set_targets")

    def set_references(self, ptrs: list):
        arr = (C.c_void_p * len(ptrs))(*ptrs)
        self._check(lib().twdic_batch_set_references(self.h, arr), "set_references")

    def run(self):
        self._check(lib().twdic_batch_run(self.h), "run")

    def run_icgn(self):
        self._check(lib().twdic_batch_run_icgn(self.h), "run_icgn")

    def sync(self):
        self._check(lib().twdic_batch_sync(self.h), "sync")

    @property
    def tuning(self) -> Tuning:
        t = Tuning()
        self._check(lib().twdic_batch_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t

    def set_tuning(self, icgn_chunk=4, fftcc_variant=3, use_graph=1, icgn_waves=5, icgn_xcd_ranges=0):
        t = Tuning(icgn_chunk, fftcc_variant, use_graph, icgn_waves, icgn_xcd_ranges)
        self._check(lib().twdic_batch_set_tuning(self.h, C.byref(t)), "set_tuning")

    def autotune(self, table_path: str | None = None, reps: int = 3):
        self._check(lib().twdic_batch_autotune(self.h, table_path.encode() if table_path else None, reps), "autotune")

    def load_tuning(self, table_path: str) -> bool:
        found = C.c_int(0)
        self._check(lib().twdic_batch_load_tuning(self.h, table_path.encode(), C.byref(found)), "load_tuning")
        return bool(found.value)

    def info(self) -> BatchInfo:
        i = BatchInfo()
        self._check(lib().twdic_batch_info_get(self.h, C.byref(i)), "info")
        return i

    def record(self, slot: int):
        self._check(lib().twdic_batch_record(self.h, slot), "record")

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        self._check(lib().twdic_batch_elapsed(self.h, a, b, C.byref(ms)), "elapsed")
        return ms.value


class _Pinned:
    def __init__(self, nbytes):
        p = C.c_void_p()
        if lib().twdic_host_alloc(C.byref(p), max(1, nbytes)) != OK:
            raise TwdicError("twdic_host_alloc failed")
        self.ptr = p.value

    def __del__(self):
        if self.ptr:
            lib().twdic_host_free(self.ptr)
            self.ptr = None


class PinnedArray(np.ndarray):
    """ndarray over pinned host memory; the allocation lives as long as the
    array (views keep their base alive)."""


def pinned_empty(shape, dtype) -> np.ndarray:
    """A numpy array in pinned host memory (freed with the array)."""
    dtype = np.dtype(dtype)
    shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
    nbytes = int(np.prod(shape)) * dtype.itemsize
    buf = _Pinned(nbytes)
    raw = (C.c_char * max(1, nbytes)).from_address(buf.ptr)
    a = np.frombuffer(raw, dtype=np.uint8, count=nbytes).view(dtype).reshape(shape).view(PinnedArray)
    a._owner = buf
    return a


class Server:
    """A twdic_server: continuous-batching serving loop over resident
    references and POI sets.  Host arrays handed to submit() are kept alive
    until the request has been waited for."""

    def __init__(self, slots=8, queue_capacity=256, hbm_budget=0, max_width=2048, max_height=2048,
                 max_pois=20000, device=0):
        cfg = ServerConfig(device, slots, queue_capacity, hbm_budget, max_width, max_height, max_pois)
        h = C.c_void_p()
        st = lib().twdic_server_create(C.byref(cfg), C.byref(h))
        if st != OK:
            raise TwdicError(f"twdic_server_create failed with status {st}")
        self.h = h.value
        self._live = {}
        self._poi_n = []

    def _check(self, st, what=""):
        if st != OK:
            raise TwdicError(f"{what}: status {st}: {lib().twdic_server_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            lib().twdic_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_reference(self, ref_id: int, img: np.ndarray):
        img = np.ascontiguousarray(img, np.uint8)
        self._check(lib().twdic_server_set_reference(self.h, ref_id, _ptr(img), img.shape[1], img.shape[0],
                                                     img.shape[1]), "set_reference")

    def add_poiset(self, poi_yx: np.ndarray) -> int:
        poi = np.ascontiguousarray(poi_yx, np.int32)
        i = C.c_int()
        self._check(lib().twdic_server_add_poiset(self.h, _ptr(poi), len(poi), C.byref(i)), "add_poiset")
        self._poi_n.append(len(poi))
        return i.value

    @staticmethod
    def outputs(n: int, pinned: bool = False) -> dict:
        mk = (lambda shp, dt: pinned_empty(shp, dt)) if pinned else (lambda shp, dt: np.zeros(shp, dt))
        return dict(u=mk(n, np.float32), v=mk(n, np.float32), zncc=mk(n, np.float32), p6=mk((n, 6), np.float32),
                    iters=mk(n, np.int32), flags=mk(n, np.uint8))

    def submit(self, ref_id: int, poiset: int, target, subset: int, fft_n: int, max_iter=20,
               delta_p=0.001, drop_if_full=False, out: dict | None = None, memory: int = MEM_PAGEABLE,
               stride: int | None = None):
        """Returns (ticket, outputs dict) or (None, None) when dropped.  `out`
        (from outputs(n)) is reused when given.  With memory=MEM_DEVICE the
        target and outputs are DeviceArrays (stride = target row bytes)."""
        n = self._poi_n[poiset]
        if out is None:
            out = self.outputs(n)
        if isinstance(target, DeviceArray):
            tp, st_ = target.ptr, stride if stride is not None else target.shape[-1]
            keep = target
        else:
            keep = np.ascontiguousarray(target, np.uint8) if memory == MEM_PAGEABLE else target
            tp, st_ = _ptr(keep), keep.shape[1]
        pt = {k: (a.ptr if isinstance(a, DeviceArray) else _ptr(a)) for k, a in out.items()}
        r = Request(ref_id, poiset, tp, st_, subset, fft_n, max_iter, delta_p, pt["u"], pt["v"], pt["zncc"],
                    pt["p6"], pt["iters"], pt["flags"], memory)
        t = C.c_uint64()
        st = lib().twdic_server_submit(self.h, C.byref(r), 1 if drop_if_full else 0, C.byref(t))
        if st == EBUSY:
            return None, None
        self._check(st, "submit")
        self._live[t.value] = (keep, out)
        return t.value, out

    def poll(self, ticket: int) -> int:
        s = C.c_int()
        self._check(lib().twdic_server_poll(self.h, ticket, C.byref(s)), "poll")
        if s.value != 0:
            self._live.pop(ticket, None)
        return s.value

    def wait(self, ticket: int):
        st = lib().twdic_server_wait(self.h, ticket)
        self._live.pop(ticket, None)
        self._check(st, "wait")

    def drain(self):
        self._check(lib().twdic_server_drain(self.h), "drain")

    def stats(self) -> ServerStats:
        s = ServerStats()
        self._check(lib().twdic_server_stats_get(self.h, C.byref(s)), "stats")
        return s

    def stats_reset(self):
        self._check(lib().twdic_server_stats_reset(self.h), "stats_reset")
