"""Helpers shared by the GPU parity tests: run the HIP path through the C ABI
on given inputs, and audit IC-GN results against an oracle run."""
import numpy as np

import oracle
from tw_amd import api


def ctx_for(case, **kw):
    ctx = api.Context(case.config(**kw))
    ctx.set_reference(case.ref)
    ctx.set_target(case.tar)
    return ctx


def run_fftcc(ctx, poi):
    n = len(poi)
    dpoi = ctx.array(poi.shape, np.int32).upload(poi)
    u = ctx.array(n, np.float32).fill(0)
    v = ctx.array(n, np.float32).fill(0)
    z = ctx.array(n, np.float32).fill(0)
    st = ctx.array(n, np.uint8).fill(255)
    ctx.fftcc(dpoi, n, u, v, z, st)
    ctx.sync()
    return dict(u=u.download(), v=v.download(), zncc=z.download(), status=st.download())


def run_icgn(ctx, poi, u0, v0):
    n = len(poi)
    dpoi = ctx.array(poi.shape, np.int32).upload(poi)
    u = ctx.array(n, np.float32).upload(u0.astype(np.float32))
    v = ctx.array(n, np.float32).upload(v0.astype(np.float32))
    p6 = ctx.array((n, 6), np.float32).fill(0)
    it = ctx.array(n, np.int32).upload(np.full(n, -1, np.int32))
    fl = ctx.array(n, np.uint8).fill(255)
    dn = ctx.array(n, np.float32).fill(0)
    ctx.icgn(dpoi, n, u, v, p6, it, fl, dn)
    ctx.sync()
    return dict(u=u.download(), v=v.download(), p6=p6.download(), iters=it.download(), flags=fl.download(),
                dpnorm=dn.download())


def check_icgn(g, o, delta_p=0.001, utol=1e-4, gtol=1e-6, band=0.05, max_flips=None):
    """Parity with the oracle o.  Iteration counts may differ only where a
    convergence decision sat inside the FP32 noise band of the loop test: the
    side that stopped first saw |dp| within `band` (relative) of delta_p.  The
    FP32 CPU path carries ~1 % rounding error in |dp| at that point (measured
    against the FP64 oracle), so this is the reference's own noise.  At most
    max_flips such POIs (default max(2, 0.2 % of the POIs))."""
    assert np.array_equal(g["flags"], o["flags"]), np.nonzero(g["flags"] != o["flags"])
    same = g["iters"] == o["iters"]
    flipped = np.nonzero(~same)[0]
    for i in flipped:
        if g["iters"][i] < o["iters"][i] and g.get("dpnorm") is not None:
            stopper = g["dpnorm"][i]
        elif g["iters"][i] > o["iters"][i]:
            stopper = o["dpnorm"][i]
        else:
            continue        # GPU stopped first but its |dp| was not returned
        assert abs(stopper / delta_p - 1.0) < band, (i, g["iters"][i], o["iters"][i], stopper)
    cap = max(2, int(0.002 * len(same))) if max_flips is None else max_flips
    assert len(flipped) <= cap, f"{len(flipped)} iteration-count flips, at most {cap} allowed"
    ok = same & (o["flags"] == oracle.SUCCESS)
    du = max(np.max(np.abs(g["u"][ok] - o["u"][ok])), np.max(np.abs(g["v"][ok] - o["v"][ok])))
    dg = np.max(np.abs(g["p6"][ok][:, [1, 2, 4, 5]] - o["p6"][ok][:, [1, 2, 4, 5]]))
    print(f"  max |du|,|dv| = {du:.3g} (bar {utol:g}); max |d grad| = {dg:.3g} (bar {gtol:g}); "
          f"iteration flips {len(flipped)}")
    assert du <= utol and dg <= gtol
    return len(flipped)
