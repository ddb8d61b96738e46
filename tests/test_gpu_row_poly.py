"""The fused IC-GN's per-target row-polynomial plane follows the target.

The single context builds the plane when the target changes and reuses it
while the target stays; once the target buffer has been handed out
(twdic_get_buffer, as cuICGN2D's drop-in GPUHandle members do) it rebuilds it
on every call, since the caller may then write the target directly.  Each
case here must give, bit for bit, what a fresh context on the final target
gives; and the fused mode must agree with the LUT mode (the reference's
materialised 16-coefficient LUT, gathered per pixel) to rounding."""
import ctypes as C

import numpy as np
import pytest

from tests import cases
from tests.gpu_helpers import ctx_for, run_icgn
from tw_amd import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    c = cases.speckle_512()
    poi = c.poi()
    rng = np.random.default_rng(77)
    u0 = np.round(rng.uniform(1.5, 3.0, len(poi))).astype(np.float32)
    v0 = np.round(rng.uniform(-2.5, -1.0, len(poi))).astype(np.float32)
    # a second target: the same reference under a different shift
    _, tar2 = api.synth_pair(512, 512, 4000, 1001, api.resample(-1.4, 0.8), noise_sigma=2.0)
    return c, poi, u0, v0, tar2


def _same(a, b):
    for k in ("u", "v", "p6", "iters", "flags"):
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def _fresh(c, tar, poi, u0, v0):
    with ctx_for(c) as ctx:
        ctx.set_target(tar)
        return run_icgn(ctx, poi, u0, v0)


def test_new_target_rebuilds_the_plane(pair):
    c, poi, u0, v0, tar2 = pair
    print("[gpu-test] row polynomials: set_target after a call", flush=True)
    with ctx_for(c) as ctx:
        first = run_icgn(ctx, poi, u0, v0)
        again = run_icgn(ctx, poi, u0, v0)          # same target: the plane is reused
        ctx.set_target(tar2)
        second = run_icgn(ctx, poi, u0, v0)
    _same(first, again)
    _same(first, _fresh(c, c.tar, poi, u0, v0))
    _same(second, _fresh(c, tar2, poi, u0, v0))


def test_target_written_through_its_buffer(pair):
    c, poi, u0, v0, tar2 = pair
    print("[gpu-test] row polynomials: target written through the handed-out buffer", flush=True)
    with ctx_for(c) as ctx:
        run_icgn(ctx, poi, u0, v0)
        ptr, nbytes = ctx.buffer(api.BUF_TAR)
        assert nbytes >= tar2.nbytes
        src = np.ascontiguousarray(tar2)
        # the context's device image has row stride = width here (512 bytes)
        rc = api.lib().twdic_memcpy(ctx.h, C.c_void_p(ptr), src.ctypes.data_as(C.c_void_p), src.nbytes, 0)
        assert rc == 0
        written = run_icgn(ctx, poi, u0, v0)
        again = run_icgn(ctx, poi, u0, v0)
    _same(written, _fresh(c, tar2, poi, u0, v0))
    _same(written, again)


def test_fused_agrees_with_the_lut_mode(pair):
    c, poi, u0, v0, _ = pair
    print("[gpu-test] row polynomials against the LUT mode", flush=True)
    fused = _fresh(c, c.tar, poi, u0, v0)
    with api.Context(c.config(interp=api.INTERP_LUT)) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        lut = run_icgn(ctx, poi, u0, v0)
    ok = (fused["flags"] == 0) & (lut["flags"] == 0)
    assert ok.mean() > 0.99 and np.array_equal(fused["flags"], lut["flags"])
    du = np.maximum(np.abs(fused["u"] - lut["u"]), np.abs(fused["v"] - lut["v"]))[ok]
    # both are the same bicubic polynomial; the iterations stop within the
    # FP32 noise of the loop test, so a few POIs may take one step more
    same_iters = (fused["iters"] == lut["iters"])[ok]
    assert same_iters.mean() > 0.99
    assert np.max(du[same_iters]) < 1e-4, np.max(du[same_iters])
