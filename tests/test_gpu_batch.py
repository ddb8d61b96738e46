"""Batched form on the GPU: image pairs of uneven sizes, subset radii and
FFT-CC window classes (radix-4 engine N = 32/64, direct-DFT engine N = 40/24)
in one set of persistent launches.  Every per-POI result must be
bit-identical to the single-context path (POIs never interact), alone or in
the mix, with or without hipGraph replay, before and after autotuning."""
import numpy as np
import pytest

from tw_amd import api

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="batched execution is switched off: its first GPU run stalled (no output "
                                      "for 180 s) and no cause has been shown")]

SPECS = [  # (w, h, S, grid, N, deformation seed)
    (256, 256, 10, 7, 32, 1),
    (384, 320, 15, 9, 64, 2),
    (512, 512, 20, 11, 40, 3),
    (300, 200, 12, 6, 24, 4),
    (448, 448, 20, 13, 64, 5),
    (512, 480, 30, 14, 64, 6),
    (320, 288, 18, 8, 56, 7),
]


def _make(spec):
    w, h, S, g, N, seed = spec
    rng = np.random.default_rng(seed)
    d = api.affine(*rng.uniform(-4, 4, 2), *rng.uniform(-0.01, 0.01, 3), w / 2, h / 2)
    ref, tar = api.synth_pair(w, h, int(0.015 * w * h), 500 + seed, d)
    m = max(N // 2, S + 6)
    ys = np.arange(m, h - m, g)
    xs = np.arange(m, w - m, g)
    yy, xx = np.meshgrid(ys, xs, indexing="ij")
    poi = np.stack([yy.ravel(), xx.ravel()], 1).astype(np.int32)
    return dict(w=w, h=h, S=S, N=N, ref=ref, tar=tar, poi=poi)


@pytest.fixture(scope="module")
def data():
    return [_make(s) for s in SPECS]


def _single(p):
    cfg = api.make_config(p["w"], p["h"], 1, 1, p["w"] - 2, p["h"] - 2, p["S"], p["S"], 1, 1, 0, 0, p["N"])
    n = len(p["poi"])
    with api.Context(cfg) as ctx:
        ctx.set_reference(p["ref"])
        ctx.set_target(p["tar"])
        dpoi = ctx.array(p["poi"].shape, np.int32).upload(p["poi"])
        u, v, z = (ctx.array(n, np.float32).fill(0) for _ in range(3))
        p6 = ctx.array((n, 6), np.float32).fill(0)
        it = ctx.array(n, np.int32).fill(0)
        fl = ctx.array(n, np.uint8).fill(255)
        ctx.fftcc(dpoi, n, u, v, z)
        ctx.icgn(dpoi, n, u, v, p6, it, fl)
        ctx.sync()
        return dict(u=u.download(), v=v.download(), zncc=z.download(), p6=p6.download(), iters=it.download(),
                    flags=fl.download())


class _Buffers:
    """Device buffers of a set of pairs, allocated through a helper context."""

    def __init__(self, data):
        self.alloc = api.Context(api.make_config(64, 64, 1, 1, 62, 62, 2, 2, 4, 4, 0, 0, 8))
        self.bufs = []
        for p in data:
            n = len(p["poi"])
            b = dict(ref=self.alloc.array((p["h"], p["w"]), np.uint8).upload(p["ref"]),
                     tar=self.alloc.array((p["h"], p["w"]), np.uint8).upload(p["tar"]),
                     poi=self.alloc.array(p["poi"].shape, np.int32).upload(p["poi"]),
                     u=self.alloc.array(n, np.float32), v=self.alloc.array(n, np.float32),
                     zncc=self.alloc.array(n, np.float32), p6=self.alloc.array((n, 6), np.float32),
                     iters=self.alloc.array(n, np.int32), flags=self.alloc.array(n, np.uint8))
            self.bufs.append(b)

    def pairs(self, data, idx=None):
        out = []
        for i in (range(len(data)) if idx is None else idx):
            p, b = data[i], self.bufs[i]
            for k in ("u", "v", "zncc", "p6", "iters", "flags"):
                b[k].fill(0)
            out.append(api.Pair(b["ref"].ptr, b["tar"].ptr, p["w"], p["w"], p["h"], 1, 1, p["w"] - 2, p["h"] - 2,
                                p["S"], p["S"], p["N"], 20, 0.001, b["poi"].ptr, len(p["poi"]), b["u"].ptr,
                                b["v"].ptr, b["zncc"].ptr, b["p6"].ptr, b["iters"].ptr, b["flags"].ptr))
        self.alloc.sync()          # the fills run on the helper context's stream
        return out

    def results(self, i):
        b = self.bufs[i]
        return {k: b[k].download() for k in ("u", "v", "zncc", "p6", "iters", "flags")}


def _same(a, b):
    for k in ("u", "v", "zncc", "p6", "iters", "flags"):
        assert np.array_equal(np.asarray(a[k]).view(np.uint8), np.asarray(b[k]).view(np.uint8)), k


def test_batch_matches_single_context_bitwise(data):
    singles = [_single(p) for p in data]
    bufs = _Buffers(data)
    with api.Batch() as bt:
        bt.set_pairs(bufs.pairs(data))
        info = bt.info()
        assert info.total_pois == sum(len(p["poi"]) for p in data)
        assert list(info.fftcc_items) == [sum(len(p["poi"]) for p in data if p["N"] == 64),
                                          sum(len(p["poi"]) for p in data if p["N"] == 32),
                                          sum(len(p["poi"]) for p in data if p["N"] not in (32, 64))]
        for graph in (1, 0, 1):                    # graph replays with eager runs in between
            bt.set_tuning(icgn_chunk=3, fftcc_variant=3, use_graph=graph)
            bt.set_pairs(bufs.pairs(data))
            bt.run()
            bt.sync()
            bt.run()                               # replay: counters are reset by a kernel node
            bt.sync()
            for i in range(len(data)):
                _same(bufs.results(i), singles[i])
    # the flags must be meaningful (mostly converged) for the comparison to bite
    assert np.mean(np.concatenate([s["flags"] for s in singles]) == 0) > 0.95
    bufs.alloc.close()


def test_alone_equals_in_mix_and_new_targets(data):
    bufs = _Buffers(data)
    with api.Batch() as bt:
        bt.set_pairs(bufs.pairs(data, [2]))
        bt.run()
        bt.sync()
        alone = bufs.results(2)
        bt.set_pairs(bufs.pairs(data))
        bt.run()
        bt.sync()
        _same(bufs.results(2), alone)
        # a new target frame for pair 0 (reference state kept resident)
        w, h, S, g, N, seed = SPECS[0]
        _, t2 = api.synth_pair(w, h, int(0.015 * w * h), 500 + seed, api.affine(-2.5, 1.25, 0.004, 0, 0, w / 2, h / 2),
                               want_ref=False)
        dt2 = bufs.alloc.array((h, w), np.uint8).upload(t2)
        bt.set_targets([dt2.ptr] + [None] * (len(data) - 1))
        bt.run()
        bt.sync()
        ref0 = _single(dict(data[0], tar=t2))
        _same(bufs.results(0), ref0)
    bufs.alloc.close()


def test_autotune_table_roundtrip(data, tmp_path):
    bufs = _Buffers(data)
    table = tmp_path / "tune.tsv"
    with api.Batch() as bt:
        bt.set_pairs(bufs.pairs(data))
        bt.autotune(str(table), reps=1)
        chosen = bt.tuning
        lines = table.read_text().strip().splitlines()
        assert len(lines) == 1 and f"pairs={len(SPECS)}" in lines[0]
        bt.run()
        bt.sync()
        first = [bufs.results(i) for i in range(len(data))]
    with api.Batch() as bt2:
        bt2.set_pairs(bufs.pairs(data))
        assert bt2.load_tuning(str(table))
        t = bt2.tuning
        assert (t.icgn_chunk, t.fftcc_variant, t.use_graph) == (chosen.icgn_chunk, chosen.fftcc_variant,
                                                                chosen.use_graph)
        bt2.run()
        bt2.sync()
        for i in range(len(data)):
            _same(bufs.results(i), first[i])
    bufs.alloc.close()
