"""Edge cases of the path on the GPU (run with -m gpu).

* Empty POI sets: FFT-CC, IC-GN and the frame step return without touching
  any output; a batch whose pairs hold no POIs runs as a no-op, and an empty
  pair inside a mix leaves its outputs untouched.  The reference rejects an
  empty set in its constructors (TW_paDIC_FFTCC2D.cpp:88-94), so these are
  the calls after construction.
* Too small buffers: the Python layer refuses a POI count its buffers cannot
  hold, on the host, before any launch.
* Tiny and ragged POI counts (1, 2, 3, 5, 7: the tails of IC-GN's four-POI
  workgroups): POIs never interact, so each POI's results are bit-identical
  to the same POI inside the full grid.  The full grid itself is held to the
  oracle by the parity tests.
"""
import numpy as np
import pytest

from tests import cases
from tests.gpu_helpers import ctx_for
from tests.test_gpu_batch import SPECS, _Buffers, _make, _same, _single
from tw_amd import api

pytestmark = pytest.mark.gpu

KEYS = ("u", "v", "zncc", "p6", "iters", "flags")


@pytest.fixture(scope="module")
def c512():
    return cases.speckle_512()


def _outputs(ctx, n):
    return dict(u=ctx.array(n, np.float32).upload(np.full(n, 7.5, np.float32)),
                v=ctx.array(n, np.float32).upload(np.full(n, -3.25, np.float32)),
                zncc=ctx.array(n, np.float32).upload(np.full(n, 2.0, np.float32)),
                p6=ctx.array((n, 6), np.float32).upload(np.full((n, 6), 9.0, np.float32)),
                iters=ctx.array(n, np.int32).upload(np.full(n, -7, np.int32)),
                flags=ctx.array(n, np.uint8).upload(np.full(n, 201, np.uint8)))


def _frame(ctx, poi):
    n = len(poi)
    dpoi = ctx.array(poi.shape, np.int32).upload(poi)
    o = _outputs(ctx, n)
    ctx.frame(dpoi, n, o["u"], o["v"], o["zncc"], o["p6"], o["iters"], o["flags"])
    ctx.sync()
    return {k: o[k].download() for k in KEYS}


def test_empty_poi_set_is_a_noop(c512):
    print("[gpu-test] empty POI sets: fftcc, icgn, frame, frame_capture", flush=True)
    poi = c512.poi()[:1]
    with ctx_for(c512) as ctx:
        dpoi = ctx.array(poi.shape, np.int32).upload(poi)
        o = _outputs(ctx, 1)
        before = {k: o[k].download() for k in KEYS}
        ctx.fftcc(dpoi, 0, o["u"], o["v"], o["zncc"])
        ctx.icgn(dpoi, 0, o["u"], o["v"], o["p6"], o["iters"], o["flags"])
        ctx.frame(dpoi, 0, o["u"], o["v"], o["zncc"], o["p6"], o["iters"], o["flags"])
        ctx.sync()
        _same({k: o[k].download() for k in KEYS}, before)
        with pytest.raises(api.TwdicError):           # a graph of nothing is refused
            ctx.frame_capture(dpoi, 0, o["u"], o["v"], o["zncc"], o["p6"], o["iters"], o["flags"])


def test_poi_count_beyond_buffers_is_refused_on_the_host(c512):
    print("[gpu-test] oversized POI counts refused before launch", flush=True)
    poi = c512.poi()[:4]
    with ctx_for(c512) as ctx:
        dpoi = ctx.array(poi.shape, np.int32).upload(poi)
        o = _outputs(ctx, 4)
        small = ctx.array(3, np.float32)
        with pytest.raises(ValueError):
            ctx.fftcc(dpoi, 5, o["u"], o["v"], o["zncc"])               # poi holds 4
        with pytest.raises(ValueError):
            ctx.icgn(dpoi, 4, small, o["v"], o["p6"], o["iters"], o["flags"])
        with pytest.raises(ValueError):
            ctx.frame(dpoi, 4, o["u"], o["v"], o["zncc"], ctx.array((3, 6), np.float32), o["iters"], o["flags"])
        with pytest.raises(ValueError):
            ctx.frame(dpoi, -1, o["u"], o["v"], o["zncc"])
        ctx.sync()


def test_tiny_and_ragged_counts_equal_the_full_grid(c512):
    poi = c512.poi()
    rng = np.random.default_rng(5)
    with ctx_for(c512) as ctx:
        full = _frame(ctx, poi)
        for n in (1, 2, 3, 5, 7):
            print(f"[gpu-test] {n} POIs against the full grid", flush=True)
            idx = np.sort(rng.choice(len(poi), n, replace=False))
            sub = _frame(ctx, np.ascontiguousarray(poi[idx]))
            _same(sub, {k: full[k][idx] for k in KEYS})
    assert np.mean(full["flags"] == 0) > 0.95


@pytest.fixture(params=["flat", "direct"])
def form(request, monkeypatch):
    monkeypatch.setenv("TWDIC_BATCH_FORM", request.param)
    return request.param


def test_batch_with_empty_and_tiny_pairs(form):
    print(f"[gpu-test] batch form {form}: empty, 3-POI and full pairs", flush=True)
    big = _make(SPECS[4])                                  # 448^2, S = 20, N = 64
    small = _make(SPECS[0])                                # 256^2, S = 10, N = 32
    tiny = dict(small, poi=np.ascontiguousarray(small["poi"][[0, 9, 23]]))
    data = [small, tiny, big]
    singles = {1: _single(tiny), 2: _single(big)}
    bufs = _Buffers(data)
    with api.Batch() as bt:
        pairs = bufs.pairs(data)
        pairs[0].n = 0                                     # pair 0 holds no POIs
        bt.set_pairs(pairs[:1])                            # nothing to do at all
        bt.run()
        bt.sync()
        bt.set_pairs(pairs)
        assert bt.info().total_pois == len(tiny["poi"]) + len(big["poi"])
        for _ in range(2):
            bt.run()
            bt.sync()
            for i in (1, 2):
                _same(bufs.results(i), singles[i])
            untouched = bufs.results(0)
            for k in KEYS:
                assert not np.any(np.asarray(untouched[k]).view(np.uint8)), k
    bufs.alloc.close()
