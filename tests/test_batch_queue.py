"""The batched form's work lists and the persistent kernels' queue loops,
emulated on the CPU with the lists of the 64-pair configuration.

The host builds, per image pair i with n_i POIs (batch_api.cpp set_pairs):
  * items: (i, k) for every POI, pair after pair (IC-GN and per-reference
    state, persistent form);
  * fitems[c]: the same per FFT-CC window class c = 0 (N = 64), 1 (N = 32),
    2 (other N) (FFT-CC, persistent form, two items per workgroup grab);
  * mitems[f]: per matrix-core kernel form (FFT-CC, flat form);
  * wg[t + 2 w]: (i, k) for k = 0, G, 2 G, ... of each pair, t = 0 when the
    subset fits the step table, w = 0 with G = 8 POIs per workgroup (one per
    half-wavefront) for pairs of at least kIcgnHalfWaveMin POIs, else w = 1
    with G = 4 (IC-GN, flat form).
and each persistent loop is
  ref state:  i = atomicAdd(c, 1)     per wavefront; exit when i >= total
  IC-GN:      b = atomicAdd(c, chunk) per wavefront; items [b, min(b + chunk, total));
              exit when b >= total
  FFT-CC:     k = atomicAdd(c, 1)     per workgroup; items 2k and 2k + 1 (the second
              only when 2k + 1 < total); exit when 2k >= total
with the counter cleared before each launch.  The emulation runs every loop
with many actors under seeded random interleavings and checks that every
actor exits after a bounded number of grabs, that every item is processed
exactly once and that no read leaves its list; the lists are checked entry by
entry against the pairs' POI counts.  The POI counts come from the same random
stream as the GPU test of this configuration (tests/cases.py batch_layout)."""
import re
from pathlib import Path

import numpy as np
import pytest

from tests import cases

CSRC = Path(__file__).resolve().parents[1] / "tw_amd" / "csrc"
# POIs per IC-GN workgroup item, as the library defines it: one per group of
# icgn_lanes_for(n) lanes of a 256-thread workgroup
_HDR = (CSRC / "twdic_internal.h").read_text()
assert "inline int icgn_lanes_for(int n) { return n >= kIcgnHalfWaveMin ? 32 : 64; }" in _HDR
HALF_WAVE_MIN = int(re.search(r"constexpr int kIcgnHalfWaveMin = (\d+);", _HDR).group(1))


def pois_per_wg(n):
    return 256 // (32 if n >= HALF_WAVE_MIN else 64)


def mma_form(N):
    return 0 if N == 64 else 1 if N > 48 else 2 if N > 32 else 3


def fclass(N):
    return 0 if N == 64 else 1 if N == 32 else 2


def work_lists(layout):
    counts = np.array([len(L["poi"]) for L in layout], np.int64)
    Ns = [L["N"] for L in layout]
    side = [2 * L["S"] + 1 for L in layout]
    pair = np.repeat(np.arange(len(layout)), counts)
    poi = np.concatenate([np.arange(n) for n in counts])
    items = np.stack([pair, poi], 1)
    fitems = [items[np.isin(pair, [i for i, N in enumerate(Ns) if fclass(N) == c])] for c in range(3)]
    mitems = [items[np.isin(pair, [i for i, N in enumerate(Ns) if mma_form(N) == f])] for f in range(4)]
    wg = [[], [], [], []]
    for i, n in enumerate(counts):
        t = (0 if side[i] * side[i] <= 2048 and side[i] <= 48 else 1) + (0 if pois_per_wg(n) == 8 else 2)
        wg[t].extend((i, k) for k in range(0, n, pois_per_wg(n)))
    wg = [np.array(w, np.int64).reshape(-1, 2) for w in wg]
    return counts, items, fitems, mitems, wg


@pytest.fixture(scope="module")
def c3():
    layout = cases.batch_layout()
    return layout, work_lists(layout)


def check_entries(lst, counts, step=1):
    assert lst.ndim == 2 and lst.shape[1] == 2
    if not len(lst):
        return
    assert lst[:, 0].min() >= 0 and lst[:, 0].max() < len(counts)
    assert np.all(lst[:, 1] >= 0) and np.all(lst[:, 1] < counts[lst[:, 0]])
    assert np.all(lst[:, 1] % step == 0)


def test_lists_name_only_existing_pois(c3):
    layout, (counts, items, fitems, mitems, wg) = c3
    assert counts.sum() == 10_553_540 and len(items) == counts.sum()
    for lst in [items, *fitems, *mitems]:
        check_entries(lst, counts)
    for t, w in enumerate(wg):
        check_entries(w, counts, 8 if t < 2 else 4)
    # every POI of every pair is in exactly one class / form list and in one
    # workgroup item
    assert sum(len(f) for f in fitems) == len(items) and sum(len(m) for m in mitems) == len(items)
    assert sum(len(w) for w in wg) == int(sum((n + pois_per_wg(n) - 1) // pois_per_wg(n) for n in counts))
    assert len(wg[0]) + len(wg[1]) > 0 and len(wg[2]) + len(wg[3]) > 0       # both widths occur in the mix


def emulate(total, actors, grab, cover, rng, max_steps):
    """Discrete-event run of a persistent loop: each step one live actor (chosen
    at random) takes its next counter grab; grab(counter) -> (new counter, base);
    cover(base) -> the item range processed, or None when the actor exits."""
    counter = 0
    live = list(range(actors))
    grabs = np.zeros(actors, np.int64)
    seen = np.zeros(total, np.int32)
    steps = 0
    while live:
        steps += 1
        assert steps <= max_steps, "a persistent loop did not drain"
        a = live[rng.integers(len(live))]
        counter, base = grab(counter)
        grabs[a] += 1
        r = cover(base)
        if r is None:
            live.remove(a)
            continue
        lo, hi = r
        assert 0 <= lo < hi <= total               # every read stays inside the list
        seen[lo:hi] += 1
    assert np.all(seen == 1)
    return grabs


@pytest.mark.parametrize("seed", [0, 1])
def test_persistent_loops_drain_and_cover_every_item_once(c3, seed):
    layout, (counts, items, fitems, mitems, wg) = c3
    rng = np.random.default_rng(seed)
    # the event-level run takes a random set of pairs holding about 60k items
    # (the whole list is 10.5M); the loops' logic depends on the total only
    # through the exit test
    order = rng.permutation(64)
    sel = order[np.cumsum(counts[order]) <= 60_000]
    if not len(sel):
        sel = order[:1]
    total = int(counts[sel].sum())
    actors = 1024                                   # 256 workgroups x 4 wavefronts
    # per-reference state: one item per grab
    g = emulate(total, actors, lambda c: (c + 1, c), lambda b: None if b >= total else (b, b + 1), rng,
                total + actors + 1)
    assert g.sum() == total + actors
    # IC-GN: chunk items per grab, every chunk the autotuner may pick
    for chunk in (1, 3, 16):
        g = emulate(total, actors, lambda c, k=chunk: (c + k, c),
                    lambda b, k=chunk: None if b >= total else (b, min(b + k, total)), rng,
                    total // chunk + actors + 2)
        assert g.sum() == -(-total // chunk) + actors
    # FFT-CC: two items per workgroup grab, per window class
    fc = [int(sum(counts[i] for i in sel if fclass(layout[i]["N"]) == c)) for c in range(3)]
    for c in range(3):
        t = fc[c]
        if not t:
            continue
        g = emulate(t, 256, lambda x: (x + 1, x), lambda k, t=t: None if 2 * k >= t else (2 * k, min(2 * k + 2, t)),
                    rng, t // 2 + 256 + 2)
        assert g.sum() == -(-t // 2) + 256


def test_counter_reset_covers_every_persistent_counter():
    """Every counter a persistent launch grabs from is cleared on the stream
    before that launch (read from the host source: the counter enum, the
    clears and the counters each launch is given).  A counter left at its
    previous launch's end would make every grab fail at once: the launch would
    skip its items (no hang), so all of them must be cleared."""
    src = (CSRC / "batch_api.cpp").read_text()
    enum = re.search(r"enum \{ (CNT_STATE = 0[^}]*)\}", src).group(1)
    cnt = {k.strip(): int(v) for k, v in (e.split("=") for e in enum.split(","))}
    cleared = set()
    for base, n in re.findall(r"launch_fill_zero\(b->d_counters \+ (CNT_\w+), (\d+)", src):
        cleared |= set(range(cnt[base], cnt[base] + int(n)))
    used = set()
    for base, off in re.findall(r"b->d_counters \+ (CNT_\w+)( \+ c)?,", src):
        used |= set(range(cnt[base], cnt[base] + (3 if off else 1)))
    assert used and used <= cleared and max(cleared) < cnt["NCNT"], (used, cleared)
