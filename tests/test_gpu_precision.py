"""Precision study on the 4096^2 sinusoidal workload with low-texture boxes:
the HIP path (FP32 warp and interpolation, FP64 sums and Cholesky solve) and
the reference's FP32 CPU path, each against the full-FP64 oracle, per-POI
u/v error percentiles over textured and low-texture subsets.  The GPU runs
all 1,018,081 POIs; the oracles run every POI whose subset touches a
low-texture box plus 20,000 others."""
import json
import os

import numpy as np
import pytest

import oracle
from tests import cases
from tw_amd import api

pytestmark = pytest.mark.gpu


def _pct(x):
    x = np.asarray(x, np.float64)
    if x.size == 0:
        return None
    return {"n": int(x.size), "p50": float(np.percentile(x, 50)), "p99": float(np.percentile(x, 99)),
            "p99.9": float(np.percentile(x, 99.9)), "max": float(x.max())}


def test_precision_study():
    case, boxes = cases.precision_4096()
    cfg = case.config()
    with api.Context(cfg) as ctx:
        ctx.set_reference(case.ref)
        ctx.set_target(case.tar)
        dpoi = ctx.poi_grid()
        n = ctx.nx * ctx.ny
        assert n == 1009 * 1009
        u, v, z = (ctx.array(n, np.float32).fill(0) for _ in range(3))
        p6 = ctx.array((n, 6), np.float32).fill(0)
        it = ctx.array(n, np.int32).fill(0)
        fl = ctx.array(n, np.uint8).fill(255)
        ctx.fftcc(dpoi, n, u, v, z)
        ctx.sync()
        iu, iv = u.download(), v.download()
        ctx.icgn(dpoi, n, u, v, p6, it, fl)
        ctx.sync()
        poi = dpoi.download()
        g = dict(u=u.download(), v=v.download(), p6=p6.download(), iters=it.download(), flags=fl.download())
    # POIs whose 41x41 subset overlaps a box (grown by the FFT window)
    near = np.zeros(n, bool)
    for x0, y0, x1, y1 in boxes:
        near |= (poi[:, 1] >= x0 - 32) & (poi[:, 1] < x1 + 32) & (poi[:, 0] >= y0 - 32) & (poi[:, 0] < y1 + 32)
    inside = np.zeros(n, bool)
    for x0, y0, x1, y1 in boxes:
        inside |= (poi[:, 1] - 20 >= x0) & (poi[:, 1] + 20 < x1) & (poi[:, 0] - 20 >= y0) & (poi[:, 0] + 20 < y1)
    rng = np.random.default_rng(4)
    sample = np.union1d(np.nonzero(near)[0], rng.choice(np.nonzero(~near)[0], 20000, replace=False))
    sp = np.ascontiguousarray(poi[sample])
    u0, v0 = iu[sample], iv[sample]
    o32 = oracle.icgn(case.ref, case.tar, case.roi, 20, 20, sp, u0, v0)
    o64 = oracle.icgn(case.ref, case.tar, case.roi, 20, 20, sp, u0.astype(np.float64), v0.astype(np.float64),
                      precision="f64")
    gs = {k: a[sample] for k, a in g.items()}
    ok = (o64["flags"] == 0) & (gs["flags"] == 0) & (o32["flags"] == 0)
    same = gs["iters"] == o32["iters"]
    tex = ok & ~near[sample]
    low = ok & inside[sample]
    edge = ok & near[sample] & ~inside[sample]

    def err(a, b, m):
        return np.maximum(np.abs(a["u"][m] - b["u"][m]), np.abs(a["v"][m] - b["v"][m]))
    tr = np.c_[0.5 * np.sin(2 * np.pi * poi[:, 1] / 256.0), 0.3 * np.cos(2 * np.pi * poi[:, 0] / 192.0)]
    gok = (g["flags"] == 0) & ~near
    report = {
        "pois_gpu": n, "pois_oracle": int(len(sample)),
        "flags_gpu": np.bincount(g["flags"], minlength=5).tolist(),
        "flags_equal_fp64_oracle": float(np.mean(gs["flags"] == o64["flags"])),
        "gpu_vs_fp64": {"textured": _pct(err(gs, o64, tex)), "box_edge": _pct(err(gs, o64, edge)),
                        "low_texture": _pct(err(gs, o64, low))},
        "fp32_path_vs_fp64": {"textured": _pct(err(o32, o64, tex)), "box_edge": _pct(err(o32, o64, edge)),
                              "low_texture": _pct(err(o32, o64, low))},
        "gpu_vs_fp32_path": {"textured": _pct(err(gs, o32, tex)), "box_edge": _pct(err(gs, o32, edge)),
                             "low_texture": _pct(err(gs, o32, low))},
        "gpu_vs_fp32_path_same_iterations": {"textured": _pct(err(gs, o32, tex & same)),
                                             "box_edge": _pct(err(gs, o32, edge & same))},
        "iteration_count_differs_textured": float(np.mean(~same[tex])),
        "gpu_vs_analytic_textured": _pct(np.maximum(np.abs(g["u"][gok] - tr[gok, 0]),
                                                    np.abs(g["v"][gok] - tr[gok, 1]))),
    }
    print("\nprecision study:", json.dumps(report, indent=1))
    out = os.environ.get("TWDIC_REPORT_DIR")
    if out:
        with open(os.path.join(out, "precision4096.json"), "w") as f:
            json.dump(report, f, indent=1)
    # Both paths evaluate warped coordinates in FP32 as the reference does; at
    # x, y in [2048, 4096) one ulp is 2.4e-4 px, which sets their common
    # distance from the FP64 oracle (p50 ~6e-5 px on textured subsets).
    # Textured subsets: the HIP path within 1e-4 px of the reference's FP32
    # path at p99 wherever both stopped after the same number of iterations,
    # and within 5e-4 px at the maximum: on this curved field the first-order
    # warp converges linearly, each iterate still moving ~1e-3 weighted px when
    # |dp| < 1e-3 stops it, so FP64 versus FP32 sums shift the stopping point
    # within that step (the FP32 path is itself 2e-4 px from FP64 there).
    # Within 0.01 px of the FP64 oracle everywhere.
    assert report["gpu_vs_fp32_path_same_iterations"]["textured"]["p99"] < 1e-4
    assert report["gpu_vs_fp32_path_same_iterations"]["textured"]["max"] < 5e-4
    assert report["iteration_count_differs_textured"] < 0.01
    assert report["gpu_vs_fp64"]["textured"]["max"] < 0.01
    # everywhere the HIP path is at least as close to FP64 as the reference's
    # own FP32 path (median and p99 of each class); on low-texture subsets
    # its FP64 sums and solve make it closer
    for cls in ("textured", "box_edge", "low_texture"):
        a, b = report["gpu_vs_fp64"][cls], report["fp32_path_vs_fp64"][cls]
        if a is None or b is None:
            continue
        assert a["p50"] <= b["p50"] * 1.05 + 1e-7, cls
        assert a["p99"] <= b["p99"] * 1.05 + 1e-7, cls
