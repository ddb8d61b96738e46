"""twdic_icgn_noalloc, the IC-GN entry that allocates nothing (per-POI
reference state in the library's fixed per-GPU scratch array), against
twdic_icgn on the same inputs through the C ABI: bit-identical u, v, p6,
iteration counts and flags, on two calls in a row, the second with a smaller
and different POI set (the scratch keeps the first call's contents); and a
POI count above the scratch's capacity is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from tests import cases
from tw_amd import api

pytestmark = pytest.mark.gpu

OUTS = ("u", "v", "p6", "iters", "flags")


def _run(ctx, poi_d, n, u0, v0, noalloc):
    u = ctx.array(n, np.float32).upload(u0)
    v = ctx.array(n, np.float32).upload(v0)
    p6 = ctx.array((n, 6), np.float32).fill(0)
    it = ctx.array(n, np.int32).fill(0)
    fl = ctx.array(n, np.uint8).fill(255)
    (ctx.icgn_noalloc if noalloc else ctx.icgn)(poi_d, n, u, v, p6, it, fl)
    ctx.sync()
    return dict(u=u.download(), v=v.download(), p6=p6.download(), iters=it.download(), flags=fl.download())


@pytest.mark.parametrize("interp", [api.INTERP_FUSED, api.INTERP_LUT])
def test_noalloc_equals_icgn_on_two_calls(interp):
    c = cases.speckle_512()
    poi = c.poi()
    print(f"[gpu-test] icgn_noalloc interp={interp}: {len(poi)} then 1000 POIs", flush=True)
    rng = np.random.default_rng(77)
    sets = [poi, np.ascontiguousarray(poi[np.sort(rng.choice(len(poi), 1000, replace=False))])]
    with api.Context(c.config(interp=interp)) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        for k, pts in enumerate(sets):
            n = len(pts)
            poi_d = ctx.array(pts.shape, np.int32).upload(pts)
            u0 = np.full(n, 2.0, np.float32)
            v0 = np.full(n, -2.0, np.float32)
            a = _run(ctx, poi_d, n, u0, v0, noalloc=True)
            b = _run(ctx, poi_d, n, u0, v0, noalloc=False)
            for key in OUTS:
                assert np.array_equal(a[key].view(np.uint8), b[key].view(np.uint8)), (k, key)
            assert np.mean(a["flags"] == 0) > 0.95, k


def test_noalloc_refuses_more_than_its_capacity():
    c = cases.speckle_512()
    with api.Context(c.config()) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        small = ctx.array(16, np.int32)
        st = api.lib().twdic_icgn_noalloc(ctx.h, C.c_void_p(small.ptr), 65537, C.c_void_p(small.ptr),
                                          C.c_void_p(small.ptr), None, None, None)
        assert st == api.EINVAL
        assert b"65536" in api.lib().twdic_last_error(ctx.h)
