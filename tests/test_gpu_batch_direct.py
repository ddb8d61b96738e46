"""Batched form in its two executions: flat (the default: one launch per
stage and kernel form over the items of all pairs) and direct
(TWDIC_BATCH_FORM=direct: the single-context kernels pair by pair).  Both run
the single-context per-POI code, so every per-POI result must be
bit-identical to a single context on the same pair, alone or in any mix,
across new targets, new references and repeated runs.  The pairs mix image
sizes, subset radii (the largest beyond the IC-GN step table) and FFT-CC
windows (32, 64 and the reference's own 2S sizes 24 and 40, which cover three
of the four matrix-core kernel forms)."""
import numpy as np
import pytest

from tests.test_gpu_batch import SPECS, _Buffers, _make, _same, _single
from tw_amd import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    return [_make(s) for s in SPECS]


@pytest.fixture(params=["flat", "direct"])
def form(request, monkeypatch):
    monkeypatch.setenv("TWDIC_BATCH_FORM", request.param)
    return request.param


def test_batch_matches_single_context_bitwise(data, form):
    print(f"[gpu-test] batch form {form}", flush=True)
    singles = [_single(p) for p in data]
    bufs = _Buffers(data)
    with api.Batch() as bt:
        bt.set_pairs(bufs.pairs(data))
        info = bt.info()
        assert info.total_pois == sum(len(p["poi"]) for p in data)
        for rep in range(2):                       # the second run reuses the resident state
            bt.run()
            bt.sync()
            for i in range(len(data)):
                _same(bufs.results(i), singles[i])
        bt.run_icgn()                              # IC-GN again from the refined guesses
        bt.sync()
    assert np.mean(np.concatenate([s["flags"] for s in singles]) == 0) > 0.95
    bufs.alloc.close()


def test_alone_equals_in_mix_new_targets_and_references(data, form):
    print(f"[gpu-test] batch form {form}", flush=True)
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
        _same(bufs.results(0), _single(dict(data[0], tar=t2)))
        # a new reference for pair 1: its per-POI state is rebuilt
        p1 = data[1]
        r2 = np.ascontiguousarray(np.roll(p1["ref"], (1, -2), (0, 1)))
        dr2 = bufs.alloc.array(r2.shape, np.uint8).upload(r2)
        bufs.alloc.sync()
        bt.set_references([None, dr2.ptr] + [None] * (len(data) - 2))
        bt.run()
        bt.sync()
        _same(bufs.results(1), _single(dict(p1, ref=r2)))
    bufs.alloc.close()


def test_direct_form_refuses_autotune(monkeypatch):
    monkeypatch.setenv("TWDIC_BATCH_FORM", "direct")
    with api.Batch() as bt:
        with pytest.raises(api.TwdicError):
            bt.autotune(None, reps=1)


def test_flat_autotune_picks_a_candidate_and_keeps_results(data, monkeypatch, tmp_path):
    """The flat form's autotuner times the four IC-GN configurations (5 or 6
    waves per SIMD, items dealt out or in per-XCD ranges), keeps one of them,
    and writes it to the table, which a new batch on the same pairs loads;
    results after tuning are still bit-identical to single contexts."""
    monkeypatch.setenv("TWDIC_BATCH_FORM", "flat")
    print("[gpu-test] flat form autotune", flush=True)
    singles = [_single(p) for p in data]
    bufs = _Buffers(data)
    table = tmp_path / "tuning.tsv"
    with api.Batch() as bt:
        bt.set_pairs(bufs.pairs(data))
        bt.autotune(str(table), reps=1)
        t = bt.tuning
        assert t.icgn_waves in (5, 6) and t.icgn_xcd_ranges in (0, 1)
        bt.run()
        bt.sync()
        for i in range(len(data)):
            _same(bufs.results(i), singles[i])
    line = table.read_text().strip().splitlines()[-1]
    fields = line.split("\t")[1].split()
    assert [int(x) for x in fields[3:5]] == [t.icgn_waves, t.icgn_xcd_ranges]
    with api.Batch() as bt2:
        bt2.set_pairs(bufs.pairs(data))
        assert bt2.load_tuning(str(table))
        t2 = bt2.tuning
        assert (t2.icgn_waves, t2.icgn_xcd_ranges) == (t.icgn_waves, t.icgn_xcd_ranges)
    bufs.alloc.close()


def test_flat_graph_replays_equal_eager_steps(data, monkeypatch):
    """The flat form's step replayed from its hipGraph (the default tuning)
    equals the eagerly launched step, bit for bit, on every replay: three
    steps over alternating targets, with eager single-context work and syncs
    between them."""
    monkeypatch.setenv("TWDIC_BATCH_FORM", "flat")
    print("[gpu-test] flat form, graph replay against eager steps", flush=True)
    w, h, S, g, N, seed = SPECS[0]
    _, t2 = api.synth_pair(w, h, int(0.015 * w * h), 700 + seed, api.affine(1.5, -0.75, 0, 0.003, 0, w / 2, h / 2),
                           want_ref=False)
    runs = {}
    for graph in (0, 1):
        bufs = _Buffers(data)
        dt2 = bufs.alloc.array((h, w), np.uint8).upload(t2)
        bufs.alloc.sync()
        out = []
        with api.Batch() as bt:
            bt.set_tuning(use_graph=graph)
            bt.set_pairs(bufs.pairs(data))
            for k in range(3):
                tar0 = dt2.ptr if k == 1 else bufs.bufs[0]["tar"].ptr
                bt.set_targets([tar0] + [None] * (len(data) - 1))
                bt.run()
                bt.sync()
                out.append([bufs.results(i) for i in range(len(data))])
                _single(data[k])                    # eager work between the steps
        runs[graph] = out
        bufs.alloc.close()
    for k in range(3):
        for i in range(len(data)):
            _same(runs[1][k][i], runs[0][k][i])
