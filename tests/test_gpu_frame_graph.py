"""The graph-captured frame step (twdic_frame_capture / twdic_frame_replay)
must give twdic_frame's results, bit for bit, on every replay: over
consecutive targets (one of them T = R), across a new reference (the replay
rebuilds the per-POI state before launching), with eager work and syncs
between replays, in the fused and LUT interpolation modes.  A replay after
the context reallocated its per-POI state is refused."""
import numpy as np
import pytest

from tests import cases
from tw_amd import api

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "zncc", "p6", "iters", "flags")


def _buffers(ctx, n):
    u, v, z = (ctx.array(n, np.float32).fill(0) for _ in range(3))
    p6 = ctx.array((n, 6), np.float32).fill(0)
    it = ctx.array(n, np.int32).fill(0)
    fl = ctx.array(n, np.uint8).fill(255)
    return u, v, z, p6, it, fl


def _frames(c, interp, graph):
    """Frames over (target, reference) steps; the fourth step has a new reference."""
    other = np.ascontiguousarray(np.roll(c.ref, (1, -2), (0, 1)))
    steps = [(c.tar, None), (c.ref, None), (c.tar, other), (c.tar, None)]
    out = []
    with api.Context(c.config(interp=interp)) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        poi = ctx.poi_grid()
        n = ctx.nx * ctx.ny
        bufs = _buffers(ctx, n)
        if graph:
            ctx.frame_capture(poi, n, *bufs)
            side = _buffers(ctx, n)
        for k, (tar, new_ref) in enumerate(steps):
            if new_ref is not None:
                ctx.set_reference(new_ref)
            ctx.set_target(tar)
            if graph:
                ctx.frame_replay()
                ctx.sync()
                # eager work between replays, on other buffers
                ctx.fftcc(poi, n, side[0], side[1], side[2])
                ctx.icgn(poi, n, side[0], side[1], side[3], side[4], side[5])
                ctx.sync()
            else:
                ctx.frame(poi, n, *bufs)
                ctx.sync()
            out.append([a.download() for a in bufs])
    return out


@pytest.mark.parametrize("interp", [api.INTERP_FUSED, api.INTERP_LUT])
def test_replayed_frame_equals_eager_frame(interp):
    c = cases.speckle_512()
    print(f"[gpu-test] frame graph replay, interp {interp}", flush=True)
    eager = _frames(c, interp, graph=False)
    replay = _frames(c, interp, graph=True)
    for f, (a, b) in enumerate(zip(eager, replay)):
        for name, x, y in zip(NAMES, a, b):
            assert np.array_equal(x, y), (f, name)
    assert np.all(replay[1][0] == 0) and np.all(replay[1][1] == 0)    # T = R


def test_replay_refused_after_state_reallocation():
    c = cases.speckle_512()
    with api.Context(c.config()) as ctx:
        ctx.set_reference(c.ref)
        ctx.set_target(c.tar)
        poi = ctx.poi_grid()
        n = ctx.nx * ctx.ny
        bufs = _buffers(ctx, n)
        ctx.frame_capture(poi, n, *bufs)
        ctx.frame_replay()
        ctx.sync()
        # a larger POI set makes the context reallocate its per-POI state
        big = np.concatenate([poi.download().reshape(-1, 2)] * 2).astype(np.int32)
        dpoi = ctx.array(big.shape, np.int32).upload(big)
        bb = _buffers(ctx, 2 * n)
        ctx.icgn(dpoi, 2 * n, bb[0], bb[1], bb[3], bb[4], bb[5])
        ctx.sync()
        with pytest.raises(api.TwdicError):
            ctx.frame_replay()
        # capturing again makes the step usable
        ctx.frame_capture(poi, n, *bufs)
        ctx.frame_replay()
        ctx.sync()
