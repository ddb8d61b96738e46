"""IC-GN's two lane layouts, chosen by the POI set's size: sets of 32,768
POIs or more run one POI per 32-lane half-wavefront, smaller sets one per
wavefront.  For each interpolation mode (fused, LUT, B-spline) the 2048^2
workload's POIs (157,609; 156,816 in the B-spline mode, whose LUT needs a
2-pixel frame right of and below the ROI) run with half-wavefronts, and a
20,000-POI subset of
the same POIs, from the same integer guesses, with whole wavefronts.  The
two differ only in the order of their exact or FP64 per-lane partial sums,
so on the subset the flags must be equal, the iteration counts equal on at
least 99.9 % of the POIs, and u, v and the gradients equal to 1e-5 px and
1e-7 where the counts agree.  A seeded 256-POI sample of the large set is
also held to the FP32 and FP64 oracles in that mode, with the bars of the
2048^2 parity test."""
import numpy as np
import pytest

import oracle
from tests import cases
from tests.gpu_helpers import check_icgn
from tw_amd import api

pytestmark = pytest.mark.gpu


def _run(ctx, poi, u0, v0):
    n = len(poi)
    dpoi = ctx.array(poi.shape, np.int32).upload(np.ascontiguousarray(poi))
    u = ctx.array(n, np.float32).upload(u0)
    v = ctx.array(n, np.float32).upload(v0)
    p6 = ctx.array((n, 6), np.float32).fill(0)
    it = ctx.array(n, np.int32).fill(0)
    fl = ctx.array(n, np.uint8).fill(255)
    dn = ctx.array(n, np.float32).fill(0)
    ctx.icgn(dpoi, n, u, v, p6, it, fl, dn)
    ctx.sync()
    return dict(u=u.download(), v=v.download(), p6=p6.download(), iters=it.download(), flags=fl.download(),
                dpnorm=dn.download())


@pytest.mark.parametrize("interp", [api.INTERP_FUSED, api.INTERP_LUT, api.INTERP_BSPLINE])
def test_half_and_whole_wavefront_layouts_agree(interp):
    c = cases.speckle_2048()
    print(f"[gpu-test] IC-GN lane layouts, 2048^2, interp {interp}", flush=True)
    roi = c.roi if interp != api.INTERP_BSPLINE else (1, 1, c.img - 3, c.img - 3)
    cfg = api.make_config(c.img, c.img, *roi, c.S, c.S, c.grid, c.grid, c.margin, c.margin, c.N, interp=interp)
    with api.Context(cfg) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        dpoi = ctx.poi_grid()
        n = ctx.nx * ctx.ny
        u, v, z = (ctx.array(n, np.float32).fill(0) for _ in range(3))
        ctx.fftcc(dpoi, n, u, v, z)
        ctx.sync()
        poi, u0, v0 = dpoi.download(), u.download(), v.download()
        big = _run(ctx, poi, u0, v0)                                   # 157,609 POIs: halves
        sub = np.sort(np.random.default_rng(7).choice(n, 20000, replace=False))
        small = _run(ctx, poi[sub], u0[sub], v0[sub])                  # 20,000 POIs: whole wavefronts
    assert n >= 32768
    b = {k: a[sub] for k, a in big.items()}
    assert np.array_equal(b["flags"], small["flags"])
    same = b["iters"] == small["iters"]
    assert same.mean() >= 0.999, same.mean()
    du = np.maximum(np.abs(b["u"] - small["u"]), np.abs(b["v"] - small["v"]))[same]
    dg = np.abs(b["p6"][:, [1, 2, 4, 5]] - small["p6"][:, [1, 2, 4, 5]])[same]
    assert du.max() <= 1e-5 and dg.max() <= 1e-7, (du.max(), dg.max())
    # the large set's results against the oracles in this mode
    mode = "bspline" if interp == api.INTERP_BSPLINE else "bicubic"
    idx = np.sort(np.random.default_rng(2049).choice(n, 256, replace=False))
    sp = np.ascontiguousarray(poi[idx])
    o = oracle.icgn(c.ref, c.tar, roi, c.S, c.S, sp, u0[idx], v0[idx], interp=mode)
    o64 = oracle.icgn(c.ref, c.tar, roi, c.S, c.S, sp, u0[idx].astype(np.float64), v0[idx].astype(np.float64),
                      precision="f64", interp=mode)
    g = {k: a[idx] for k, a in big.items()}
    # the bars of the 2048^2 parity test: relative to the FP32 path's own
    # distance from FP64 (values, iteration flips and their |dp| band)
    ok = (o["iters"] == o64["iters"]) & (o["flags"] == 0)
    own_u = float(np.max(np.maximum(np.abs(o["u"] - o64["u"]), np.abs(o["v"] - o64["v"]))[ok]))
    own_g = float(np.max(np.abs(o["p6"][:, [1, 2, 4, 5]] - o64["p6"][:, [1, 2, 4, 5]])[ok]))
    own = int(np.count_nonzero(o["iters"] != o64["iters"]))
    cap = max(2, int(0.002 * len(idx)), 3 * own + 2)
    fl = np.nonzero(o["iters"] != o64["iters"])[0]
    stop = np.where(o["iters"][fl] < o64["iters"][fl], o["dpnorm"][fl], o64["dpnorm"][fl])
    band = max(0.05, 1.25 * float(np.max(np.abs(stop / 1e-3 - 1.0)))) if len(fl) else 0.05
    print(f"  FP32 oracle vs FP64: u, v {own_u:.3g} px, gradients {own_g:.3g}, flips {own}", flush=True)
    check_icgn(g, o64, utol=max(5e-5, 1.25 * own_u), gtol=max(2e-6, 1.5 * own_g), band=band, max_flips=cap)
    check_icgn(g, o, utol=max(1e-4, 2 * own_u), gtol=max(3e-6, 2 * own_g), band=band, max_flips=cap)
