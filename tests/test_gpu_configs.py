"""GPU tests on the benchmark configurations themselves, through the C ABI:
plain_1k (the reference's own CPU-runnable case), batched_512 at its full
size with 500-step graph replays, spiking_20k (spikes driven through the
gap junctions), and the driver's file inputs against the reference's
golden traces.

Tolerances (max |dV_axon| in mV) are stated at each assertion; spike
agreement is stated as matched fractions and step offsets.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import infoli_amd as I
import oracle_lib as O
from infoli_amd import analysis as A
from infoli_amd import workloads as W
from test_gpu_parity import PRECS, make_net, oracle_ref, say

pytestmark = pytest.mark.gpu


def load(net, wl, gcal=None, iapp=None):
    W.load_into(net, wl)
    if gcal is not None:
        net.set_gcal(gcal)
    if iapp is not None:
        net.set_iapp(iapp)
    return net


# ---- plain_1k: 1,000 cells, G(n, 0.1), w = 0.04, g_CaL 0.68, pulse on [10000, 20000) ----

@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 2e-4), ("fp32k", 2e-4)])
def test_plain_1k_first_100ms_vs_oracle(prec, tol):
    """The whole network, the first 2,000 steps (100 ms), against the oracle
    in the same arithmetic (FP64 to 1e-9; FP32 to 2e-4, the v_exp_f32 and
    fused multiply-add of the gap term against expf) and against FP64
    (1e-3, the precision budget)."""
    say(f"plain_1k first 100 ms {prec}")
    wl = W.plain_1k()
    net = load(I.Network(wl["n"], precision=prec), wl)
    e = net.exec
    out = net.advance(2000, trace_stride=10)
    net.close()
    ref, _ = oracle_ref(prec, wl["n"], 2000, wl["row_ptr"], wl["col"], wl["w"], wl["gcal"], amp=wl["amp"],
                        t0=wl["t0"], t1=wl["t1"], stride=10)
    err = float(np.abs(out - ref).max())
    say(f"  exec {e}; max |dV| vs {prec} oracle {err:.3e} mV")
    assert err < tol
    if prec != "fp64":
        ref64, _ = oracle_ref("fp64", wl["n"], 2000, wl["row_ptr"], wl["col"], wl["w"], wl["gcal"], amp=wl["amp"],
                              t0=wl["t0"], t1=wl["t1"], stride=10)
        assert np.abs(out - ref64).max() < 1e-3


def test_plain_1k_full_run():
    """All 20,000 steps (1 s) with the 6 uA/cm^2 pulse of 500-1000 ms.
    Every cell starts alike with the same g_CaL and stimulus, so the network
    stays synchronous bit for bit (I_c = 0 exactly) in each arithmetic; the
    FP64 trajectory is the single-cell FP64 oracle's (1e-6 over the second,
    the pulse's spike included); FP32 with compensated sums fires the FP64
    run's spikes within one step."""
    say("plain_1k full 20,000 steps")
    wl = W.plain_1k()
    nets = {p: load(I.Network(wl["n"], precision=p), wl) for p in ("fp64", "fp32k")}
    cell0 = {p: [] for p in nets}
    for t in range(0, wl["steps"], 2000):
        for p, net in nets.items():
            tr = net.advance(2000, trace_stride=1)
            assert np.all(tr == tr[:, :1]), f"{p}: cells differ at steps [{t}, {t + 2000})"
            cell0[p].append(tr[:, 0])
    for net in nets.values():
        net.close()
    a64 = np.concatenate(cell0["fp64"])
    a32 = np.concatenate(cell0["fp32k"])
    st = O.initial_state(1)
    st[O.G_CAL] = 0.68
    ref, _ = O.run(O.FP64, st, wl["steps"], amp=wl["amp"], t0=wl["t0"], t1=wl["t1"])
    err = float(np.abs(a64 - ref[:, 0]).max())
    c64, s64 = A.crossings(a64[:, None], None, -20.0, 0)
    c32, s32 = A.crossings(a32[:, None], None, -20.0, 0)
    agree = A.spike_agreement(c32, s32, c64, s64, window=50)
    say(f"  FP64 vs single-cell oracle {err:.3e} mV; spikes {agree}")
    assert err < 1e-6
    assert agree["spikes_b"] >= 1 and agree["matched"] == agree["spikes_b"] == agree["spikes_a"]
    assert agree["max_abs_dt_steps"] <= 1


# ---- batched_512: 512 networks of 64-4,096 cells, 500-step graph replays ----

@pytest.fixture(scope="module")
def batched():
    wl = W.batched_512()
    n = wl["n"]
    # per-cell g_CaL and Iapp instead of one per network: the cells of a
    # network then differ and the gap junctions carry current
    gcal = I.gen_uniform(W.MASTER_SEED, 12, n, 0.5, 1.1)
    iapp = I.gen_uniform(W.MASTER_SEED, 13, n, -1.0, 3.0)
    sizes = wl["sizes"]
    order = np.argsort(sizes, kind="stable")
    rng = np.random.default_rng(7)
    sample = sorted(set(order[:3]) | set(order[-3:]) | set(rng.choice(len(sizes), 10, replace=False)))
    return wl, gcal, iapp, [int(k) for k in sample]


def sub_network(wl, k):
    offs, rp, col, w = wl["partition"], wl["row_ptr"], wl["col"], wl["w"]
    a, b = int(offs[k]), int(offs[k + 1])
    e0, e1 = int(rp[a]), int(rp[b])
    return a, b, rp[a:b + 1] - e0, col[e0:e1] - a, w[e0:e1]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32k", 5e-4)])
def test_batched_512_heterogeneous_vs_oracle(batched, prec, tol):
    """The full batch (485,518 cells, 97.6M entries) in one context under the
    heuristic's strategy (tiled, one tile per block staging its networks'
    senders), 1,000 steps as two 500-step graph replays; 16 of the networks
    (the three smallest, the three largest, ten others) against the oracle
    run on each network alone: FP64 to 1e-9, FP32 compensated to 5e-4 (as
    the other FP32 oracle comparisons with spikes)."""
    wl, gcal, iapp, sample = batched
    say(f"batched_512 {prec}: {wl['n']} cells, networks {sample}")
    net = load(I.Network(wl["n"], precision=prec), wl, gcal, iapp)
    net.set_exec(steps_per_launch=500)
    e = net.exec
    assert e["gather"] == I.GATHER_TILED and e["mode"] == I.EXEC_GRAPH
    net.prepare(1000, 5)
    out = net.advance(1000, trace_stride=5)
    launches = net.last_launches()
    net.close()
    say(f"  exec {e}, launches {launches}")
    assert launches == {"graph_replays": 2, "direct_steps": 0, "captures": 0}
    worst = 0.0
    spiking = 0
    for k in sample:
        a, b, rp, col, w = sub_network(wl, k)
        ref, _ = oracle_ref(prec, b - a, 1000, rp, col, w, gcal[a:b], iapp[a:b], stride=5)
        worst = max(worst, float(np.abs(out[:, a:b] - ref).max()))
        spiking += int((out[:, a:b] > -20).any(axis=0).sum())
    say(f"  max |dV| over the sampled networks {worst:.3e} mV; {spiking} sampled cells spike")
    assert worst < tol
    assert spiking > 0


def test_batched_512_gap_currents_nonzero(batched):
    """With per-cell parameters the networks leave the synchronous state:
    after 200 steps the gap-junction currents are far from zero."""
    wl, gcal, iapp, _ = batched
    net = load(I.Network(wl["n"], precision="fp64"), wl, gcal, iapp)
    net.advance(200)
    ic = net.gap_currents()
    net.close()
    say(f"batched_512 |I_c| after 200 steps: max {np.abs(ic).max():.3e}, mean {np.abs(ic).mean():.3e}")
    assert np.abs(ic).max() > 0.1 and (np.abs(ic) > 1e-6).mean() > 0.9


def chunk_cells(n, tsize, kb):
    """The library's sender chunk for an LDS budget (a multiple of 64 cells)."""
    cap = kb * 1024 // tsize // 64 * 64
    nch = -(-n // cap)
    return min(cap, -(-(-(-n // nch)) // 64) * 64)


def rows_in_one_tile(wl, a, b, R, C):
    """Whether every row of cells [a, b) sits in one tile of the batch's
    layout: its receiver block stages its whole sender range (at most C
    cells from a 64-aligned start), or the network lies inside one sender
    chunk [k C, (k + 1) C)."""
    rp, col, n = wl["row_ptr"], wl["col"], wl["n"]
    for blk in range(a // R, (b - 1) // R + 1):
        c0, c1 = blk * R, min(n, blk * R + R)
        snd = col[rp[c0]:rp[c1]]
        if len(snd) == 0 or snd.max() - snd.min() // 64 * 64 + 1 <= C:
            continue
        if a // C != (b - 1) // C:
            return False
    return True


def test_batched_512_fp64_equals_per_network_contexts(batched):
    """FP64: the batch in one context against each sampled network alone in
    its own context with the same strategy (same lanes per receiver).  Where
    every row of the network is summed as one tile's partial in the batch
    (as it is alone), the traces are equal bit for bit; a network whose rows
    the batch splits at a sender-chunk boundary sums them in two partials,
    one more rounding, and stays within 1e-9 mV over 500 steps."""
    wl, gcal, iapp, sample = batched
    say("batched_512 fp64 vs per-network contexts")
    net = load(I.Network(wl["n"], precision="fp64"), wl, gcal, iapp)
    net.set_exec(steps_per_launch=500)
    e = net.exec
    out = net.advance(500, trace_stride=5)
    net.close()
    C = chunk_cells(wl["n"], 8, e["lds_kb"])
    exact = 0
    for k in sample:
        a, b, rp, col, w = sub_network(wl, k)
        one = make_net(b - a, "fp64", rp, col, w, gcal[a:b], iapp[a:b])
        one.set_exec(mode=I.EXEC_GRAPH, gather=I.GATHER_TILED, lanes_per_cell=e["lanes_per_cell"],
                     tile_cells=e["tile_cells"], lds_kb=e["lds_kb"], block=e["block"], ode_split=e["ode_split"],
                     steps_per_launch=500)
        tr = one.advance(500, trace_stride=5)
        one.close()
        if rows_in_one_tile(wl, a, b, e["tile_cells"], C):
            assert np.array_equal(tr, out[:, a:b]), f"network {k} ({b - a} cells)"
            exact += 1
        else:
            assert np.abs(tr - out[:, a:b]).max() < 1e-9, f"network {k} ({b - a} cells)"
    say(f"  {exact} of {len(sample)} sampled networks bit for bit")
    assert exact >= len(sample) // 2


# ---- spiking_20k: spikes through the gap junctions ----

def test_spiking_20k_spike_times():
    """spiking_20k in FP32 with compensated sums against FP64, both on the
    GPU: max |dV| over the first 100 ms within the 1e-3 mV budget; the cells
    fire from about 100 ms on with the junctions carrying current; spike
    times in [100, 200) ms agree to one step for at least 99.5% of the FP64
    spikes, and in [200, 300) ms, as the chaotic dynamics separate the
    runs, at least 97% of them match within the 50-step window."""
    say("spiking_20k spike times")
    wl = W.spiking_20k()
    nets = {p: load(I.Network(wl["n"], precision=p), wl) for p in ("fp32k", "fp64")}
    tr = {p: [] for p in nets}
    for t in range(0, wl["steps"], 500):
        for p, net in nets.items():
            tr[p].append(net.advance(500, trace_stride=1).astype(np.float32))
        if t + 500 == 2000:
            ic = nets["fp64"].gap_currents()
    for net in nets.values():
        net.close()
    a = np.concatenate(tr["fp32k"])
    b = np.concatenate(tr["fp64"])
    dv = float(np.abs(a[:2000].astype(np.float64) - b[:2000]).max())
    windows = {}
    for lo, hi in ((2000, 4000), (4000, 6000)):
        ca, sa = A.crossings(a[lo:hi], a[lo - 1], -20.0, lo)
        cb, sb = A.crossings(b[lo:hi], b[lo - 1], -20.0, lo)
        windows[(lo, hi)] = A.spike_agreement(ca, sa, cb, sb, window=50)
    say(f"  max |dV| first 100 ms {dv:.3e} mV; |I_c| at 100 ms max {np.abs(ic).max():.3e}; windows {windows}")
    assert dv <= 1e-3
    assert np.abs(ic).max() > 0.05
    w1, w2 = windows[(2000, 4000)], windows[(4000, 6000)]
    assert w1["spikes_b"] > wl["n"]          # about three spikes per cell per 100 ms window
    assert w1["matched"] >= 0.995 * w1["spikes_b"] and w1["max_abs_dt_steps"] <= 1
    assert w2["spikes_b"] > wl["n"] and w2["matched"] >= 0.97 * w2["spikes_b"]


# ---- the driver's file inputs (-m dense matrix, -r triplets, -i Iapp file) ----

@pytest.mark.parametrize("case", ["st_single_cell", "st_reg10_iapp", "st_rand20"])
@pytest.mark.parametrize("conn_format", ["dense", "triplets"])
def test_driver_file_inputs_reproduce_st_golden(golden, case, conn_format):
    """infoli.x reads the connectivity (dense [sender][receiver] matrix,
    src/single_threaded/infoli.c:158-191, or send/rec/cond triplets,
    src/regression_test/stable_version.c:226-252) and the per-step Iapp file
    (src/utils/utils.c:70-122) that the reference's FP64 simulator was run
    on, in the text it read them from, and its FP64 trace equals the
    reference's golden trace to 1e-9 mV."""
    say(f"driver files {case} {conn_format}")
    z = golden(case)
    conn, iapp, ref = z["conn"], z["iapp"], z["trace"]
    n = conn.shape[0]
    exe = os.path.join(O.ROOT, "infoli_amd", "bin", "infoli.x")
    with tempfile.TemporaryDirectory() as d:
        if conn_format == "dense":
            with open(os.path.join(d, "conn.txt"), "w") as fh:
                for row in conn:
                    fh.write(" ".join("%.4f" % v for v in row) + "\n")
            copt = ["-m", "conn.txt"]
        else:
            with open(os.path.join(d, "conn.txt"), "w") as fh:
                for s in range(n):
                    for r in range(n):
                        if conn[s, r] != 0:
                            fh.write("%d %d %.4f\n" % (s, r, conn[s, r]))
            copt = ["-r", "conn.txt"]
        with open(os.path.join(d, "iapp.txt"), "w") as fh:
            for row in iapp:
                fh.write(" ".join("%.3f" % v for v in row) + "\n")
        cmd = [exe, "-n", str(n), *copt, "-i", "iapp.txt", "-x", "fp64", "-k", "1",
               "-c", os.path.join(O.GOLDEN, "default.conf"), "-o", "out.txt"]
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        rows = []
        for ln in open(os.path.join(d, "out.txt")).read().splitlines():
            if not ln.strip() or ln.startswith(("#", "Execution")):
                continue
            rows.append([float(x) for x in ln.split()[2:2 + n]])
    out = np.array(rows)
    assert out.shape == ref.shape
    err = float(np.abs(out - ref).max())
    say(f"  max |dV| vs the reference's golden trace {err:.3e} mV")
    assert err < 1e-9
