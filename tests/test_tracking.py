"""POI tracking around a reference reset (TW/TW_Core/cuda_utils.cu:174-239,
called at fftcctworkerthread.cpp:419-437,476-481): the oracle restatement
pinned to numpy's truncation and FP32 addition, and the HIP kernels held
bit-exact to it, including the invalidation of the per-POI reference state
when the cached POI array moves."""
import numpy as np
import pytest

import oracle
from tests import cases
from tw_amd import api


def _data(n, seed):
    rng = np.random.default_rng(seed)
    poi = rng.integers(40, 2000, (n, 2)).astype(np.int32)
    u = rng.uniform(-12, 12, n).astype(np.float32)
    v = rng.uniform(-12, 12, n).astype(np.float32)
    # exact integers, values just below/above integers and signed zeros
    u[:8] = np.array([3.0, -3.0, 2.9999998, -2.9999998, 0.0, -0.0, 0.5, -0.5], np.float32)
    v[:8] = np.array([-1.0, 1.0, -0.9999999, 0.9999999, -0.0, 0.0, -7.5, 7.5], np.float32)
    return poi, u, v


def test_oracle_tracking_closed_forms():
    poi, u, v = _data(1000, 3)
    moved = oracle.track_update_poi(poi, u, v)
    # int() truncates toward zero: numpy trunc
    assert np.array_equal(moved[:, 0], poi[:, 0] + np.trunc(v).astype(np.int32))
    assert np.array_equal(moved[:, 1], poi[:, 1] + np.trunc(u).astype(np.int32))
    assert np.array_equal(oracle.track_current_poi(poi, u, v), moved)
    assert moved[2].tolist() == [poi[2, 0], poi[2, 1] + 2] and moved[3].tolist() == [poi[3, 0], poi[3, 1] - 2]
    au, av = oracle.track_accumulate_uv(u, v, v, u)
    assert np.array_equal(au, (u + v).astype(np.float32)) and np.array_equal(av, (v + u).astype(np.float32))
    # the 7-argument form: y takes trunc(v) - trunc(accumulated u), x takes
    # trunc(u) - trunc(accumulated v); with zero accumulations it is the
    # 5-argument form, with acc = (u, v) it moves by trunc(v) - trunc(u)
    acc_u = np.random.default_rng(5).uniform(-9, 9, len(u)).astype(np.float32)
    acc_v = np.random.default_rng(6).uniform(-9, 9, len(u)).astype(np.float32)
    got = oracle.track_current_poi_acc(poi, u, v, acc_u, acc_v)
    assert np.array_equal(got[:, 0], poi[:, 0] + np.trunc(v).astype(np.int32) - np.trunc(acc_u).astype(np.int32))
    assert np.array_equal(got[:, 1], poi[:, 1] + np.trunc(u).astype(np.int32) - np.trunc(acc_v).astype(np.int32))
    z = np.zeros_like(u)
    assert np.array_equal(oracle.track_current_poi_acc(poi, u, v, z, z), moved)
    same = oracle.track_current_poi_acc(poi, u, v, u, v)
    assert np.array_equal(same[:, 0] - poi[:, 0], np.trunc(v).astype(np.int32) - np.trunc(u).astype(np.int32))


@pytest.mark.gpu
def test_gpu_tracking_bit_exact():
    poi, u, v = _data(100003, 7)
    n = len(poi)
    cfg = api.make_config(64, 64, 1, 1, 62, 62, 2, 2, 4, 4, 0, 0, 8)
    with api.Context(cfg) as ctx:
        dp = ctx.array(poi.shape, np.int32).upload(poi)
        du = ctx.array(n, np.float32).upload(u)
        dv = ctx.array(n, np.float32).upload(v)
        out = ctx.array(poi.shape, np.int32).fill(0)
        ctx.track_current_poi(dp, n, du, dv, out)
        ctx.track_update_poi(dp, n, du, dv)
        au = ctx.array(n, np.float32).upload(v)
        av = ctx.array(n, np.float32).upload(u)
        ctx.track_accumulate_uv(n, du, dv, au, av)
        ctx.sync()
        want = oracle.track_update_poi(poi, u, v)
        assert np.array_equal(out.download(), want)
        assert np.array_equal(dp.download(), want)
        wu, wv = oracle.track_accumulate_uv(v, u, u, v)
        assert np.array_equal(au.download(), wu) and np.array_equal(av.download(), wv)
        # the 7-argument form (dp has moved: it equals `want` now)
        acc_u = np.random.default_rng(8).uniform(-30, 30, n).astype(np.float32)
        acc_v = np.random.default_rng(9).uniform(-30, 30, n).astype(np.float32)
        dau = ctx.array(n, np.float32).upload(acc_u)
        dav = ctx.array(n, np.float32).upload(acc_v)
        out2 = ctx.array(poi.shape, np.int32).fill(0)
        ctx.track_current_poi_acc(dp, n, du, dv, dau, dav, out2)
        ctx.sync()
        assert np.array_equal(out2.download(), oracle.track_current_poi_acc(want, u, v, acc_u, acc_v))


@pytest.mark.gpu
def test_gpu_moving_cached_pois_rebuilds_reference_state():
    c = cases.speckle_512()
    poi = c.poi()
    n = len(poi)
    shift_u = np.full(n, 2.6, np.float32)       # POIs move by (0, +2) in x
    shift_v = np.zeros(n, np.float32)
    moved = oracle.track_update_poi(poi, shift_u, shift_v)
    f = oracle.fftcc_exact(c.ref, c.tar, moved, 32)
    u0, v0 = f["u"].astype(np.float32), f["v"].astype(np.float32)
    with api.Context(c.config()) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        dp = ctx.array(poi.shape, np.int32).upload(poi)
        u = ctx.array(n, np.float32).upload(u0)
        v = ctx.array(n, np.float32).upload(v0)
        ctx.icgn(dp, n, u, v)                    # builds the state for the original POIs
        ctx.track_update_poi(dp, n, ctx.array(n, np.float32).upload(shift_u), ctx.array(n, np.float32).upload(shift_v))
        u.upload(u0)
        v.upload(v0)
        p6 = ctx.array((n, 6), np.float32)
        it = ctx.array(n, np.int32).fill(0)
        fl = ctx.array(n, np.uint8).fill(255)
        ctx.icgn(dp, n, u, v, p6, it, fl)        # must use state rebuilt for the moved POIs
        ctx.sync()
        got = dict(u=u.download(), v=v.download(), iters=it.download(), flags=fl.download())
    with api.Context(c.config()) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        dp = ctx.array(moved.shape, np.int32).upload(moved)
        u = ctx.array(n, np.float32).upload(u0)
        v = ctx.array(n, np.float32).upload(v0)
        it = ctx.array(n, np.int32).fill(0)
        fl = ctx.array(n, np.uint8).fill(255)
        ctx.icgn(dp, n, u, v, None, it, fl)
        ctx.sync()
        fresh = dict(u=u.download(), v=v.download(), iters=it.download(), flags=fl.download())
    for k in ("u", "v", "iters", "flags"):
        assert np.array_equal(got[k], fresh[k]), k
