"""Host-side buffer checks of the C ABI: every entry that takes device
buffers measures each against the HIP allocation that holds it
(twdic_device_span) and refuses a short one with TWDIC_EBOUNDS, and a pointer
into no allocation with TWDIC_EINVAL, before launching anything.

The span query is checked on valid buffers first; the refusals are then
exercised only where that same query has already reported the buffer short,
so no kernel is ever launched on a buffer it would overrun."""
import ctypes as C

import numpy as np
import pytest

from tw_amd import api

pytestmark = pytest.mark.gpu


def test_span_query_and_short_buffers_refused():
    cfg = api.make_config(512, 512, 1, 1, 510, 510, 8, 8, 2, 2, 0, 0, 16)
    ref, tar = api.synth_pair(512, 512, 4000, 5, api.affine(1.2, -0.7))
    with api.Context(cfg) as ctx:
        ctx.set_reference(ref)
        ctx.set_target(tar)
        # the query on valid buffers
        a = ctx.array(1000, np.uint8)
        s0 = api.device_span(a.ptr)
        assert s0 >= 1000
        assert api.device_span(a.ptr + 100) == s0 - 100
        host = np.zeros(64, np.uint8)
        with pytest.raises(api.TwdicError):
            api.device_span(host.ctypes.data)
        # a POI set whose u buffer the query reports too short
        poi = ctx.poi_grid()
        n = ctx.nx * ctx.ny
        short_u = ctx.array(4, np.float32)
        assert api.device_span(short_u.ptr) < 4 * n
        v, z = ctx.array(n, np.float32), ctx.array(n, np.float32)
        L = api.lib()
        st = L.twdic_fftcc(ctx.h, C.c_void_p(poi.ptr), n, C.c_void_p(short_u.ptr), C.c_void_p(v.ptr),
                           C.c_void_p(z.ptr), None)
        assert st == api.EBOUNDS, st
        assert b"d_u" in L.twdic_last_error(ctx.h)
        p6 = ctx.array((n, 6), np.float32)
        it, fl = ctx.array(n, np.int32), ctx.array(n, np.uint8)
        st = L.twdic_icgn(ctx.h, C.c_void_p(poi.ptr), n, C.c_void_p(short_u.ptr), C.c_void_p(v.ptr),
                          C.c_void_p(p6.ptr), C.c_void_p(it.ptr), C.c_void_p(fl.ptr), None)
        assert st == api.EBOUNDS, st
        # a host pointer where a device buffer belongs
        hu = np.zeros(n, np.float32)
        st = L.twdic_fftcc(ctx.h, C.c_void_p(poi.ptr), n, C.c_void_p(hu.ctypes.data), C.c_void_p(v.ptr),
                           C.c_void_p(z.ptr), None)
        assert st == api.EINVAL, st
        # memset past the end of an allocation
        assert L.twdic_memset(ctx.h, C.c_void_p(a.ptr), 0, s0 + 1) == api.EBOUNDS
        # and the same calls with full-size buffers still run
        u = ctx.array(n, np.float32).fill(0)
        ctx.fftcc(poi, n, u, v, z)
        ctx.icgn(poi, n, u, v, p6, it, fl)
        ctx.sync()
        assert np.mean(fl.download() == 0) > 0.9
