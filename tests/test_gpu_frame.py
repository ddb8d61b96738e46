"""The frame step pipelines FFT-CC of one POI chunk with IC-GN of the
previous one on two streams.  Its results must be bit-identical to the two
stages called back to back (the same kernels over the same POIs), for any
chunk count, and across consecutive frames and a new target."""
import numpy as np
import pytest

from tests import cases
from tw_amd import api

pytestmark = pytest.mark.gpu


def _run(c, chunks, monkeypatch, split):
    monkeypatch.setenv("TWDIC_FRAME_CHUNKS", str(chunks))
    out = []
    with api.Context(c.config()) as ctx:
        ctx.set_reference(c.ref)
        poi = ctx.poi_grid()
        n = ctx.nx * ctx.ny
        u, v, z = (ctx.array(n, np.float32).fill(0) for _ in range(3))
        p6 = ctx.array((n, 6), np.float32).fill(0)
        it = ctx.array(n, np.int32).fill(0)
        fl = ctx.array(n, np.uint8).fill(255)
        for tar in (c.tar, c.ref, c.tar):           # consecutive frames, one of them T = R
            ctx.set_target(tar)
            if split:
                ctx.fftcc(poi, n, u, v, z)
                ctx.icgn(poi, n, u, v, p6, it, fl)
            else:
                ctx.frame(poi, n, u, v, z, p6, it, fl)
            ctx.sync()
            out.append([a.download() for a in (u, v, z, p6, it, fl)])
    return out


@pytest.mark.parametrize("chunks", [4, 7, 16])
def test_pipelined_frame_equals_two_stages(chunks, monkeypatch):
    c = cases.speckle_2048()
    ref = _run(c, 1, monkeypatch, split=True)
    got = _run(c, chunks, monkeypatch, split=False)
    for f, (a, b) in enumerate(zip(ref, got)):
        for name, x, y in zip(("u", "v", "zncc", "p6", "iters", "flags"), a, b):
            assert np.array_equal(x, y), (f, name)
    # the identity frame converged at zero displacement everywhere
    assert np.all(got[1][0] == 0) and np.all(got[1][5] == 0)
