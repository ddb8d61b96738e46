"""Real-image fixture: tests/golden/tw_core_pair.npz holds a 320x240 crop of
the image pair that ships with TW_Core (TW/TW_Core/1.bmp, 2.bmp; its gross
shift is (0, 0)) and the oracle's outputs on it, written by
tests/golden/make_golden.py.  The CPU test pins the oracle to the stored
vectors (a regression fixture on real imagery); the GPU tests hold the HIP
path to the parity bars on it."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from tw_amd import api

FIX = Path(__file__).resolve().parent / "golden" / "tw_core_pair.npz"


@pytest.fixture(scope="module")
def gold():
    d = np.load(FIX)              # plain arrays, no pickles
    return {k: d[k] for k in d.files}


def test_oracle_reproduces_fixture(gold):
    S, G, M, N = (int(x) for x in gold["params"])
    roi = tuple(int(x) for x in gold["roi"])
    nx, ny, ok = oracle.num_poi(roi[2], roi[3], S, S, G, G, M, M)
    poi = oracle.poi_grid(roi[0], roi[1], nx, ny, M, M, S, S, G, G)
    assert np.array_equal(poi, gold["poi"])
    ex = oracle.fftcc_exact(gold["ref"], gold["tar"], poi, N)
    assert np.array_equal(ex["u"], gold["fftcc_u"]) and np.array_equal(ex["v"], gold["fftcc_v"])
    assert np.max(np.abs(ex["zncc"] - gold["fftcc_zncc"])) <= 1e-7
    o = oracle.icgn(gold["ref"], gold["tar"], roi, S, S, poi, ex["u"].astype(np.float32), ex["v"].astype(np.float32))
    assert np.array_equal(o["flags"], gold["icgn_flags"]) and np.array_equal(o["iters"], gold["icgn_iters"])
    # the oracle is deterministic per POI: the stored FP32 results come back to 1e-6 px
    assert np.max(np.abs(o["u"] - gold["icgn_u"])) <= 1e-6 and np.max(np.abs(o["v"] - gold["icgn_v"])) <= 1e-6
    # FP32 path vs the stored FP64 oracle: within the 1e-4 px bar
    assert np.max(np.abs(gold["icgn_u"] - gold["icgn64_u"])) < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("interp", [api.INTERP_FUSED, api.INTERP_LUT])
def test_gpu_on_real_pair(gold, interp):
    from tests.gpu_helpers import check_icgn, run_fftcc, run_icgn
    S, G, M, N = (int(x) for x in gold["params"])
    roi = tuple(int(x) for x in gold["roi"])
    h, w = gold["ref"].shape
    cfg = api.make_config(w, h, *roi, S, S, G, G, M, M, N, interp=interp)
    with api.Context(cfg) as ctx:
        ctx.set_reference(gold["ref"])
        ctx.set_target(gold["tar"])
        assert np.array_equal(ctx.poi_grid().download(), gold["poi"])
        f = run_fftcc(ctx, gold["poi"])
        # integer peaks bit-exact, ZNCC within 1e-6 of the exact value
        assert np.array_equal(f["u"], gold["fftcc_u"].astype(np.float32))
        assert np.array_equal(f["v"], gold["fftcc_v"].astype(np.float32))
        assert np.max(np.abs(f["zncc"] - gold["fftcc_zncc"])) <= 1e-6
        g = run_icgn(ctx, gold["poi"], f["u"], f["v"])
    o32 = dict(u=gold["icgn_u"], v=gold["icgn_v"], p6=gold["icgn_p6"], iters=gold["icgn_iters"],
               flags=gold["icgn_flags"], dpnorm=gold["icgn_dpnorm"])
    # against the FP32 oracle: u, v within 1e-4 px, gradients within 3e-6
    check_icgn(g, o32, gtol=3e-6)
    # against the FP64 oracle: u, v within 5e-5 px, gradients within 2e-6
    o64 = dict(u=gold["icgn64_u"], v=gold["icgn64_v"], p6=gold["icgn64_p6"], iters=gold["icgn64_iters"],
               flags=gold["icgn64_flags"], dpnorm=gold["icgn64_dpnorm"])
    check_icgn(g, o64, utol=5e-5, gtol=2e-6)
