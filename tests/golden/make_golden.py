"""Writes tests/golden/tw_core_pair.npz: a crop of the real image pair that
ships with TW_Core (TW/TW_Core/1.bmp, 2.bmp, 640x480 8-bit) and the oracle's
outputs on it (exact-integer FFT-CC, FP32 and FP64 IC-GN).

Run in a checkout that has the reference next to it:
    python tests/golden/make_golden.py /root/reference
The GPU tests read only the .npz."""
import struct
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle  # noqa: E402

# crop (x0, y0, w, h) of the 640x480 frames, and the DIC set-up on it
CROP = (160, 120, 320, 240)
S, G, M, N = 15, 8, 4, 32


def read_bmp8(path: Path) -> np.ndarray:
    """8-bit paletted BMP (BITMAPINFOHEADER, uncompressed) -> grey uint8."""
    b = path.read_bytes()
    assert b[:2] == b"BM"
    off = struct.unpack_from("<I", b, 10)[0]
    w, h = struct.unpack_from("<ii", b, 18)
    bpp, comp = struct.unpack_from("<HI", b, 28)
    assert bpp == 8 and comp == 0, (bpp, comp)
    ncol = struct.unpack_from("<I", b, 46)[0] or 256
    pal = np.frombuffer(b, np.uint8, 4 * ncol, 54).reshape(ncol, 4)   # B, G, R, 0
    grey = np.round(0.299 * pal[:, 2] + 0.587 * pal[:, 1] + 0.114 * pal[:, 0]).astype(np.uint8)
    row = (w + 3) & ~3
    idx = np.frombuffer(b, np.uint8, row * abs(h), off).reshape(abs(h), row)[:, :w]
    if h > 0:
        idx = idx[::-1]                                                  # bottom-up rows
    return grey[idx]


def main(ref_root: str):
    core = Path(ref_root) / "TW" / "TW_Core"
    r, t = read_bmp8(core / "1.bmp"), read_bmp8(core / "2.bmp")
    x0, y0, w, h = CROP
    R = np.ascontiguousarray(r[y0:y0 + h, x0:x0 + w])
    T = np.ascontiguousarray(t[y0:y0 + h, x0:x0 + w])
    roi = (1, 1, w - 2, h - 2)
    nx, ny, ok = oracle.num_poi(roi[2], roi[3], S, S, G, G, M, M)
    assert ok
    poi = oracle.poi_grid(roi[0], roi[1], nx, ny, M, M, S, S, G, G)
    ex = oracle.fftcc_exact(R, T, poi, N)
    u0, v0 = ex["u"].astype(np.float32), ex["v"].astype(np.float32)
    o32 = oracle.icgn(R, T, roi, S, S, poi, u0, v0)
    o64 = oracle.icgn(R, T, roi, S, S, poi, u0, v0, precision="f64")
    out = Path(__file__).resolve().parent / "tw_core_pair.npz"
    np.savez_compressed(out, ref=R, tar=T, roi=np.array(roi, np.int32), params=np.array([S, G, M, N], np.int32),
                        poi=poi, fftcc_u=ex["u"], fftcc_v=ex["v"], fftcc_zncc=ex["zncc"], fftcc_status=ex["status"],
                        icgn_u=o32["u"], icgn_v=o32["v"], icgn_p6=o32["p6"], icgn_iters=o32["iters"],
                        icgn_flags=o32["flags"], icgn_dpnorm=o32["dpnorm"],
                        icgn64_u=o64["u"], icgn64_v=o64["v"], icgn64_p6=o64["p6"], icgn64_iters=o64["iters"],
                        icgn64_flags=o64["flags"], icgn64_dpnorm=o64["dpnorm"])
    print(out, out.stat().st_size, "bytes;", len(poi), "POIs; flags", np.bincount(o32["flags"], minlength=5),
          "mean iters", o32["iters"][o32["flags"] == 0].mean(), "u range", ex["u"].min(), ex["u"].max(),
          "v range", ex["v"].min(), ex["v"].max())


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
