"""Precision-analysis helpers (host only): threshold crossings across trace
chunks and spike-time matching."""
import numpy as np

from infoli_amd import analysis as A


def test_crossings_across_chunks_equal_whole_trace():
    rng = np.random.default_rng(1)
    steps, cells = 400, 7
    tr = np.cumsum(rng.normal(0, 8, (steps, cells)), axis=0) - 20
    c_all, s_all = A.crossings(tr, None, -20.0, 0)
    got_c, got_s, prev = [], [], None
    for t0 in range(0, steps, 37):
        c, s = A.crossings(tr[t0:t0 + 37], prev, -20.0, t0)
        got_c.append(c)
        got_s.append(s)
        prev = tr[min(t0 + 37, steps) - 1]
    c2, s2 = np.concatenate(got_c), np.concatenate(got_s)
    assert sorted(zip(c_all, s_all)) == sorted(zip(c2, s2))
    # brute force: step s is a crossing of cell c iff tr[s-1] < th <= tr[s]
    brute = [(c, s) for c in range(cells) for s in range(1, steps) if tr[s - 1, c] < -20 <= tr[s, c]]
    assert sorted(zip(c_all, s_all)) == sorted(brute)


def test_crossing_on_the_threshold_counts_once():
    tr = np.array([[-30.0], [-20.0], [-20.0], [-25.0], [10.0]])
    c, s = A.crossings(tr, None, -20.0, 100)
    assert list(s) == [101, 104] and list(c) == [0, 0]


def test_spike_agreement_counts_and_shifts():
    # cell 0: B at 100, 300; A at 102, 299 -> dt +2, -1
    # cell 1: B at 50; A none -> only_b
    # cell 2: A at 70 only -> only_a; B at 500 with A at 700 (outside window) -> both unmatched
    ca = np.array([0, 0, 2, 2]); sa = np.array([102, 299, 70, 700])
    cb = np.array([0, 0, 1, 2]); sb = np.array([100, 300, 50, 500])
    r = A.spike_agreement(ca, sa, cb, sb, window=50)
    assert r["matched"] == 2
    assert r["only_b"] == 2 and r["only_a"] == 2
    assert r["max_abs_dt_steps"] == 2
    assert abs(r["mean_abs_dt_steps"] - 1.5) < 1e-12
    assert r["exact_fraction"] == 0.0


def test_spike_agreement_identical_runs():
    rng = np.random.default_rng(2)
    c = rng.integers(0, 50, 300)
    s = rng.integers(0, 10000, 300)
    r = A.spike_agreement(c, s, c, s)
    assert r["matched"] == 300 and r["only_a"] == 0 and r["only_b"] == 0
    assert r["exact_fraction"] == 1.0 and r["max_abs_dt_steps"] == 0
