"""Precision analysis of voltage traces (host-side numpy, no GPU code).

The dynamics are chaotic, so two arithmetics (FP32 with compensated sums
against FP64) are compared the way the model is used: the largest V_axon
difference over an initial window (100 ms = 2,000 steps at dt = 0.05 ms),
then the agreement of spike times after it.  A spike is an upward crossing
of V_axon through a threshold (-20 mV by default: the axonal spike of the
inferior-olive cell overshoots to +35..+75 mV from a -60 mV rest, the
subthreshold oscillation stays below -39 mV).
"""
import numpy as np


def crossings(trace, prev, threshold, t0):
    """Upward threshold crossings in a [steps][cells] trace chunk.

    prev: the row before the chunk (None at the start).  Returns (cells,
    steps) arrays, steps absolute (t0 = step of trace[0]), ordered by step
    then cell."""
    trace = np.asarray(trace)
    if prev is not None:
        trace = np.vstack([np.asarray(prev)[None, :], trace])
        t0 = t0 - 1
    below = trace[:-1] < threshold
    above = trace[1:] >= threshold
    s, c = np.nonzero(below & above)
    return c.astype(np.int64), (s + 1 + t0).astype(np.int64)


def spike_agreement(cells_a, steps_a, cells_b, steps_b, window=50):
    """Match the spikes of run A to those of run B cell by cell: in time
    order, each spike of B takes the first unmatched spike of A within
    `window` steps.  Returns counts and the matched spike-time differences
    (A - B, in steps)."""
    order_a = np.lexsort((steps_a, cells_a))
    order_b = np.lexsort((steps_b, cells_b))
    ca, sa = np.asarray(cells_a)[order_a], np.asarray(steps_a)[order_a]
    cb, sb = np.asarray(cells_b)[order_b], np.asarray(steps_b)[order_b]
    diffs = []
    ia = 0
    only_b = 0
    na = len(ca)
    used = np.zeros(na, bool)
    for k in range(len(cb)):
        c, t = cb[k], sb[k]
        while ia < na and (ca[ia] < c or (ca[ia] == c and sa[ia] < t - window)):
            ia += 1
        j = ia
        found = -1
        while j < na and ca[j] == c and sa[j] <= t + window:
            if not used[j]:
                found = j
                break
            j += 1
        if found < 0:
            only_b += 1
        else:
            used[found] = True
            diffs.append(int(sa[found] - t))
    diffs = np.asarray(diffs, np.int64)
    return {
        "spikes_a": int(na),
        "spikes_b": int(len(cb)),
        "matched": int(len(diffs)),
        "only_a": int(na - used.sum()),
        "only_b": int(only_b),
        "max_abs_dt_steps": int(np.abs(diffs).max()) if len(diffs) else 0,
        "mean_abs_dt_steps": float(np.abs(diffs).mean()) if len(diffs) else 0.0,
        "exact_fraction": float((diffs == 0).mean()) if len(diffs) else 1.0,
    }


def compare_runs(net_a, net_b, steps, window_steps=2000, threshold=-20.0, chunk=250,
                 match_window=50, progress=None):
    """Advance two networks in lockstep for `steps` steps with a V_axon
    record every step: max |V_a - V_b| over the first `window_steps` steps,
    and the spike-time agreement of the steps after them (all steps if the
    run is not longer than the window).  progress(t, maxdv), if given, is
    called after every 2,000 steps compared."""
    n = net_a.n
    maxdv = 0.0
    prev_a = prev_b = None
    ev = {"a": ([], []), "b": ([], [])}
    t = 0
    after0 = window_steps if steps > window_steps else 0
    while t < steps:
        k = min(chunk, steps - t)
        ta = net_a.advance(k, trace_stride=1, allow_nonfinite=True)
        tb = net_b.advance(k, trace_stride=1, allow_nonfinite=True)
        if t < window_steps:
            m = min(k, window_steps - t)
            maxdv = max(maxdv, float(np.abs(ta[:m] - tb[:m]).max()))
        for key, tr, prev in (("a", ta, prev_a), ("b", tb, prev_b)):
            c, s = crossings(tr, prev, threshold, t)
            keep = s >= after0
            ev[key][0].append(c[keep])
            ev[key][1].append(s[keep])
        prev_a, prev_b = ta[-1], tb[-1]
        t += k
        if progress is not None and (t % 2000 == 0 or t == steps):
            progress(t, maxdv)
    ca, sa = (np.concatenate(x) if x else np.zeros(0, np.int64) for x in ev["a"])
    cb, sb = (np.concatenate(x) if x else np.zeros(0, np.int64) for x in ev["b"])
    out = {"cells": n, "steps": steps, "window_steps": min(window_steps, steps),
           "max_abs_dv_mV": maxdv, "spikes_from_step": after0, "threshold_mV": threshold}
    out.update(spike_agreement(ca, sa, cb, sb, window=match_window))
    return out
