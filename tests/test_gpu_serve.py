"""Serving loop (include/twdic_serve.h) on the GPU: 1024x1024 frames deformed
from resident references (strains within +-0.5 %, translations within +-4
px), POI sets of 500 to 20,000 POIs with subset radius 10, 15 or 20.  Each
request's results must be bit-identical to the single-context path, alone and
in a mix of concurrent requests, across a reference replacement, and under
memory-limited admission; dropped requests follow the bounded-queue rule."""
import numpy as np
import pytest

import oracle
from tw_amd import api

pytestmark = pytest.mark.gpu

W = H = 1024
CLASSES = [(10, 32), (15, 32), (20, 64)]       # (S, N)


def _frames(seed, count):
    rng = np.random.default_rng(seed)
    ref = None
    tars = []
    for k in range(count):
        d = api.affine(*rng.uniform(-4, 4, 2), *rng.uniform(-0.005, 0.005, 3), W / 2, H / 2)
        r, t = api.synth_pair(W, H, int(0.0153 * W * H), seed, d, noise_sigma=2.0, noise_seed=seed * 100 + k,
                              want_ref=(ref is None))
        if ref is None:
            ref = r
        tars.append(t)
    return ref, tars


def _poiset(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(40, H - 40, n), rng.integers(40, W - 40, n)], 1).astype(np.int32)


@pytest.fixture(scope="module")
def world():
    ref0, tars0 = _frames(7001, 4)
    ref1, tars1 = _frames(7002, 3)
    sets = [_poiset(n, 50 + i) for i, n in enumerate([500, 3000, 9000, 20000])]
    return dict(refs=[ref0, ref1], tars=[tars0, tars1], sets=sets)


def _single(ref, tar, poi, S, N):
    cfg = api.make_config(W, H, 1, 1, W - 2, H - 2, S, S, 1, 1, 0, 0, N)
    n = len(poi)
    with api.Context(cfg) as ctx:
        ctx.set_reference(ref)
        ctx.set_target(tar)
        dp = ctx.array(poi.shape, np.int32).upload(poi)
        u, v, z = (ctx.array(n, np.float32).fill(0) for _ in range(3))
        p6 = ctx.array((n, 6), np.float32).fill(0)
        it = ctx.array(n, np.int32).fill(0)
        fl = ctx.array(n, np.uint8).fill(0)
        ctx.fftcc(dp, n, u, v, z)
        ctx.icgn(dp, n, u, v, p6, it, fl)
        ctx.sync()
        return dict(u=u.download(), v=v.download(), zncc=z.download(), p6=p6.download(), iters=it.download(),
                    flags=fl.download())


def _same(a, b):
    for k in ("u", "v", "zncc", "p6", "iters", "flags"):
        assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), k


def test_alone_equals_mix_equals_single_context(world):
    rng = np.random.default_rng(3)
    with api.Server(slots=8, queue_capacity=512, max_width=W, max_height=H, max_pois=20000) as srv:
        srv.set_reference(0, world["refs"][0])
        srv.set_reference(1, world["refs"][1])
        ids = [srv.add_poiset(p) for p in world["sets"]]
        S, N = CLASSES[1]
        t, alone = srv.submit(0, ids[1], world["tars"][0][1], S, N)
        srv.wait(t)
        # the same request inside a mix of 60 others of every class
        tickets = []
        for k in range(60):
            r = int(rng.integers(0, 2))
            s_, n_ = CLASSES[int(rng.integers(0, 3))]
            tickets.append(srv.submit(r, ids[int(rng.integers(0, 4))], world["tars"][r][int(rng.integers(0, 3))],
                                      s_, n_)[0])
            if k == 30:
                tm, mixed = srv.submit(0, ids[1], world["tars"][0][1], S, N)
        for tk in tickets + [tm]:
            srv.wait(tk)
        st = srv.stats()
        assert st.completed == 62 and st.failed == 0 and st.dropped == 0
        assert st.slots_busy_peak > 1               # requests overlapped on the device
        assert st.state_hits > 0 and st.state_builds >= 1
    _same(mixed, alone)
    # this build serves host memory only: other memory kinds are refused
    with api.Server(slots=2, queue_capacity=8, max_width=W, max_height=H, max_pois=20000) as srv:
        srv.set_reference(0, world["refs"][0])
        pid = srv.add_poiset(world["sets"][0])
        with pytest.raises(api.TwdicError):
            srv.submit(0, pid, world["tars"][0][0], S, N, memory=api.MEM_PINNED)
    single = _single(world["refs"][0], world["tars"][0][1], world["sets"][1], S, N)
    _same(alone, single)
    # and agrees with the oracle on a sample of the POIs
    poi = world["sets"][1][:400]
    ex = oracle.fftcc_exact(world["refs"][0], world["tars"][0][1], poi, N)
    o = oracle.icgn(world["refs"][0], world["tars"][0][1], (1, 1, W - 2, H - 2), S, S, poi,
                    ex["u"].astype(np.float32), ex["v"].astype(np.float32))
    assert np.array_equal(alone["flags"][:400], o["flags"])
    ok = (o["flags"] == 0) & (alone["iters"][:400] == o["iters"])
    assert ok.mean() > 0.98
    assert np.max(np.abs(alone["u"][:400][ok] - o["u"][ok])) < 1e-4
    assert np.max(np.abs(alone["v"][:400][ok] - o["v"][ok])) < 1e-4
    assert np.max(np.abs(alone["zncc"][:400] - ex["zncc"])[ex["status"] == 0]) < 1e-6


def test_reference_replacement_in_flight(world):
    with api.Server(slots=4, queue_capacity=512, max_width=W, max_height=H, max_pois=20000) as srv:
        srv.set_reference(0, world["refs"][0])
        pid = srv.add_poiset(world["sets"][2])
        S, N = CLASSES[2]
        before = [srv.submit(0, pid, world["tars"][0][k], S, N) for k in range(3)]
        srv.set_reference(0, world["refs"][1])           # later requests use the new reference
        after = [srv.submit(0, pid, world["tars"][1][k], S, N) for k in range(3)]
        for t, _ in before + after:
            srv.wait(t)
    for k in (0, 2):
        _same(before[k][1], _single(world["refs"][0], world["tars"][0][k], world["sets"][2], S, N))
        _same(after[k][1], _single(world["refs"][1], world["tars"][1][k], world["sets"][2], S, N))


def test_memory_admission_and_dropping(world):
    # budget: one reference (1 MiB image + 4 MiB gradients), the POI set and
    # state (20,000 x 392 B) and about two requests' buffers (1 MiB frame +
    # 16 MiB of its row polynomials + 0.8 MiB outputs each)
    budget = 50 << 20
    with api.Server(slots=8, queue_capacity=8, hbm_budget=budget, max_width=W, max_height=H,
                    max_pois=20000) as srv:
        srv.set_reference(0, world["refs"][0])
        pid = srv.add_poiset(world["sets"][3])
        S, N = CLASSES[0]
        res = []
        dropped = 0
        for k in range(40):
            t, out = srv.submit(0, pid, world["tars"][0][k % 4], S, N, drop_if_full=True)
            if t is None:
                dropped += 1
            else:
                res.append((k, t, out))
        for _, t, _ in res:
            srv.wait(t)
        st = srv.stats()
        assert st.hbm_peak <= budget and st.failed == 0
        assert st.slots_busy_peak < 8                   # memory, not slots, limited admission
        assert dropped == st.dropped and dropped > 0    # a full queue drops, as Concurrent_Buffer
        assert st.completed == len(res)
    ref_out = _single(world["refs"][0], world["tars"][0][res[0][0] % 4], world["sets"][3], S, N)
    _same(res[0][2], ref_out)
    for k, _, out in res:
        if k % 4 == res[0][0] % 4:
            _same(out, ref_out)


def test_mix_sampled_against_oracle(world):
    """60 concurrent requests of every class, both references and several
    targets; a seeded sample of 24 POIs of every request is checked against
    the oracles (FFT-CC ZNCC within 1e-6 of the exact-integer value; IC-GN
    from the exact integer guesses with the bars of the 2048^2 workload test,
    scaled by the reference FP32 path's own distance from FP64 on the same
    POIs)."""
    from tests.gpu_helpers import check_icgn
    rng = np.random.default_rng(11)
    with api.Server(slots=8, queue_capacity=512, max_width=W, max_height=H, max_pois=20000) as srv:
        srv.set_reference(0, world["refs"][0])
        srv.set_reference(1, world["refs"][1])
        ids = [srv.add_poiset(p) for p in world["sets"]]
        reqs = []
        for k in range(60):
            r = int(rng.integers(0, 2))
            c = int(rng.integers(0, 3))
            si = int(rng.integers(0, 4))
            ti = int(rng.integers(0, 3))
            S, N = CLASSES[c]
            t, out = srv.submit(r, ids[si], world["tars"][r][ti], S, N)
            reqs.append((t, out, r, si, ti, c))
        for t, *_ in reqs:
            srv.wait(t)
        st = srv.stats()
        assert st.completed == 60 and st.failed == 0 and st.slots_busy_peak > 1
    # group the samples by (reference, target, class): one oracle run each
    groups = {}
    for t, out, r, si, ti, c in reqs:
        poi = world["sets"][si]
        idx = rng.choice(len(poi), 24, replace=False)
        groups.setdefault((r, ti, c), []).append((out, poi, idx))
    g_s, o_s, o64_s = [], [], []
    for (r, ti, c), members in groups.items():
        S, N = CLASSES[c]
        ref, tar = world["refs"][r], world["tars"][r][ti]
        sp = np.ascontiguousarray(np.concatenate([poi[idx] for _, poi, idx in members]))
        ex = oracle.fftcc_exact(ref, tar, sp, N)
        zn = np.concatenate([out["zncc"][idx] for out, _, idx in members])
        assert np.max(np.abs(zn - ex["zncc"])) < 1e-6, (r, ti, c)
        u0, v0 = ex["u"].astype(np.float32), ex["v"].astype(np.float32)
        o_s.append(oracle.icgn(ref, tar, (1, 1, W - 2, H - 2), S, S, sp, u0, v0))
        o64_s.append(oracle.icgn(ref, tar, (1, 1, W - 2, H - 2), S, S, sp, u0.astype(np.float64),
                                 v0.astype(np.float64), precision="f64"))
        g_s.append({k: np.concatenate([out[k][idx] for out, _, idx in members])
                    for k in ("u", "v", "p6", "iters", "flags")})

    def cat(rs):
        return {k: np.concatenate([x[k] for x in rs]) for k in rs[0] if k != "times_ms"}
    g, o, o64 = cat(g_s), cat(o_s), cat(o64_s)
    assert np.mean(g["flags"] == 0) > 0.95
    same = (o["iters"] == o64["iters"]) & (o["flags"] == 0)
    own_u = float(np.max(np.maximum(np.abs(o["u"] - o64["u"]), np.abs(o["v"] - o64["v"]))[same]))
    own_g = float(np.max(np.abs(o["p6"][:, [1, 2, 4, 5]] - o64["p6"][:, [1, 2, 4, 5]])[same]))
    own = int(np.count_nonzero(o["iters"] != o64["iters"]))
    fl = np.nonzero(o["iters"] != o64["iters"])[0]
    stop = np.where(o["iters"][fl] < o64["iters"][fl], o["dpnorm"][fl], o64["dpnorm"][fl])
    band = max(0.05, 1.25 * float(np.max(np.abs(stop / 1e-3 - 1.0)))) if len(fl) else 0.05
    cap = max(2, int(0.002 * len(g["u"])), 3 * own + 2)
    print(f"  {len(g['u'])} sampled POIs in {len(groups)} (reference, target, class) groups; FP32 oracle vs FP64: "
          f"u, v {own_u:.3g} px, gradients {own_g:.3g}, iteration flips {own}")
    check_icgn(g, o64, utol=max(5e-5, 1.25 * own_u), gtol=max(2e-6, 1.5 * own_g), band=band, max_flips=cap)
    check_icgn(g, o, utol=max(1e-4, 2 * own_u), gtol=max(3e-6, 2 * own_g), band=band, max_flips=cap)
