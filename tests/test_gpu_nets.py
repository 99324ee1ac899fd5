"""GPU tests of whole networks per workgroup (INFOLI_GATHER_NETS, the
persistent launch that advances many steps with no barrier between
workgroups): bit for bit against the direct gather at one lane per cell,
against the oracle, across launch boundaries and checkpoints, with every
stimulus form; the heuristic's choice on a sweep of small networks; and the
refusal of networks one workgroup cannot hold.
"""
import numpy as np
import pytest

import infoli_amd as I
from infoli_amd import workloads as W
from test_gpu_parity import make_net, oracle_ref, say

pytestmark = pytest.mark.gpu

NETS = dict(mode=I.EXEC_PERSISTENT, gather=I.GATHER_NETS)
DIRECT1 = dict(mode=I.EXEC_GRAPH, gather=I.GATHER_DIRECT, lanes_per_cell=1)


def uneven_batch(seed=9, n_nets=150, lo=1, hi=1500, p=0.08):
    """Independent networks of log-uniform sizes in [lo, hi] (single cells
    included), each G(N, p) with w ~ U[0.02, 0.06], side by side."""
    u = I.gen_uniform(seed, 5, n_nets, np.log(lo), np.log(hi))
    sizes = np.maximum(1, np.round(np.exp(u)).astype(np.int64))
    offs = np.zeros(n_nets + 1, np.int64)
    offs[1:] = np.cumsum(sizes)
    rps, cols, ws, base = [np.zeros(1, np.int64)], [], [], 0
    for k in range(n_nets):
        rp, col, w = I.gen_symmetric(seed * 1000 + k, int(sizes[k]), p, 0.02, 0.06)
        rps.append(rp[1:] + base)
        cols.append(col + offs[k])
        ws.append(w)
        base += int(rp[-1])
    n = int(offs[-1])
    return (n, offs, np.concatenate(rps), np.concatenate(cols).astype(np.int32), np.concatenate(ws),
            I.gen_uniform(seed, 2, n, 0.5, 1.1), I.gen_uniform(seed, 3, n, -1.0, 3.0))


def run(prec, batch, steps, stride, exec_kw, pulse=None, table=None, splits=None):
    n, offs, rp, col, w, gcal, iapp = batch
    net = make_net(n, prec, rp, col, w, gcal, iapp)
    net.set_partition(offs)
    if pulse is not None:
        net.set_iapp_pulse(*pulse)
    if table is not None:
        net.set_iapp_table(*table)
    net.set_exec(steps_per_launch=100, **exec_kw)
    e = net.exec
    trs = [net.advance(k, trace_stride=stride) for k in (splits or [steps])]
    st = net.state()
    net.close()
    return e, np.concatenate(trs), st


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32k"])
def test_nets_bitwise_equal_direct_one_lane(prec):
    """150 uneven networks (1 to 1,500 cells; 1,000 in FP64, where a
    workgroup holds 1,024), per-cell g_CaL and Iapp, a pulse to a masked
    subset and a per-step Iapp table over part of the run: the whole state
    and every trace record equal the direct gather's at one lane per cell,
    bit for bit, over 350 steps (three full 100-step launches and a partial
    one)."""
    batch = uneven_batch(hi=1000 if prec == "fp64" else 1500)
    n = batch[0]
    mask = (I.gen_uniform(4, 4, n, 0.0, 1.0) < 0.3).astype(np.uint8)
    tab = I.gen_uniform(5, 6, 40 * n, -0.5, 0.5).reshape(40, n)
    kw = dict(pulse=(4.0, 50, 200, mask), table=(120, tab))
    say(f"nets vs direct L1 {prec}: {n} cells")
    e, tr, st = run(prec, batch, 350, 7, NETS, **kw)
    assert e["gather"] == I.GATHER_NETS and e["mode"] == I.EXEC_PERSISTENT and e["lanes_per_cell"] == 1
    _, tr1, st1 = run(prec, batch, 350, 7, DIRECT1, **kw)
    assert np.array_equal(tr, tr1)
    assert np.array_equal(st, st1)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32k"])
def test_nets_compartment_threads_bitwise_equal_direct_one_lane(prec):
    """Networks of at most 300 cells (ranges of at most 320 cells run one
    thread per compartment): 300 uneven networks with a pulse mask and an
    Iapp table, 350 steps; state and trace equal the direct gather's at one
    lane per cell, bit for bit."""
    batch = uneven_batch(seed=21, n_nets=300, lo=1, hi=300, p=0.1)
    n = batch[0]
    mask = (I.gen_uniform(6, 4, n, 0.0, 1.0) < 0.4).astype(np.uint8)
    tab = I.gen_uniform(7, 6, 30 * n, -0.5, 0.5).reshape(30, n)
    kw = dict(pulse=(3.0, 40, 250, mask), table=(200, tab))
    say(f"nets (one thread per compartment) vs direct L1 {prec}: {n} cells")
    e, tr, st = run(prec, batch, 350, 5, NETS, **kw)
    assert e["gather"] == I.GATHER_NETS
    _, tr1, st1 = run(prec, batch, 350, 5, DIRECT1, **kw)
    assert np.array_equal(tr, tr1)
    assert np.array_equal(st, st1)


@pytest.mark.parametrize("prec,steps,tol", [("fp64", 300, 1e-9), ("fp32k", 60, 2e-4)])
def test_nets_vs_oracle(prec, steps, tol):
    """A batch of uneven networks against the oracle in the same arithmetic:
    FP64 to 1e-9 mV over 300 steps; FP32 compensated to 2e-4 mV over 60
    steps, as the block-diagonal default test (further on, the cells spike
    and the hardware exp2 and fused multiply-add of the FP32 gap term move
    spike onsets by a fraction of a step, mV apart during the upstroke)."""
    batch = uneven_batch(seed=11, n_nets=60, hi=900)
    n, offs, rp, col, w, gcal, iapp = batch
    say(f"nets vs oracle {prec}: {n} cells")
    _, tr, _ = run(prec, batch, steps, 10, NETS)
    ref, _ = oracle_ref(prec, n, steps, rp, col, w, gcal, iapp, stride=10)
    assert np.abs(tr - ref).max() < tol


def test_nets_split_advances_and_checkpoint_bitwise(tmp_path):
    """Advances of 150 and 151 steps (partial persistent launches) and a
    checkpoint written and loaded between them give the state and trace of
    one 301-step advance, bit for bit."""
    batch = uneven_batch(seed=13, n_nets=40, hi=700)
    _, tr, st = run("fp32k", batch, 301, 3, NETS)
    _, tr2, st2 = run("fp32k", batch, 301, 3, NETS, splits=[150, 151])
    assert np.array_equal(tr, tr2) and np.array_equal(st, st2)
    n, offs, rp, col, w, gcal, iapp = batch
    a = make_net(n, "fp32k", rp, col, w, gcal, iapp)
    a.set_partition(offs)
    a.set_exec(steps_per_launch=100, **NETS)
    t1 = a.advance(150, trace_stride=3)
    path = str(tmp_path / "nets.ck")
    a.save_checkpoint(path)
    a.close()
    b = make_net(n, "fp32k", rp, col, w, gcal, iapp)
    b.set_partition(offs)
    b.set_exec(steps_per_launch=100, **NETS)
    b.load_checkpoint(path)
    t2 = b.advance(151, trace_stride=3)
    st3 = b.state()
    b.close()
    assert np.array_equal(np.concatenate([t1, t2]), tr) and np.array_equal(st3, st)


def test_sweep_256_default_is_whole_networks_and_matches_oracle():
    """The sweep of 256 networks of 256 cells: the heuristic advances it as
    whole networks per workgroup, 500 steps per launch; with per-cell
    parameters (so that I_c != 0) 1,000 steps match the FP64 oracle on four
    of the networks run alone (FP32 compensated, 5e-4 mV)."""
    wl = W.sweep_256()
    n = wl["n"]
    gcal = I.gen_uniform(17, 2, n, 0.5, 1.1)
    iapp = I.gen_uniform(17, 3, n, -1.0, 3.0)
    net = I.Network(n, precision="fp32k")
    W.load_into(net, wl)
    net.set_gcal(gcal)
    net.set_iapp(iapp)
    e = net.exec
    say(f"sweep_256 default exec {e}")
    assert e["gather"] == I.GATHER_NETS and e["mode"] == I.EXEC_PERSISTENT and e["steps_per_launch"] == 500
    out = net.advance(1000, trace_stride=50)
    launches = net.last_launches()
    net.close()
    assert launches["graph_replays"] == 2 and launches["direct_steps"] == 0
    offs, rp, col, w = wl["partition"], wl["row_ptr"], wl["col"], wl["w"]
    for k in (0, 77, 160, 255):
        a, b = int(offs[k]), int(offs[k + 1])
        e0, e1 = int(rp[a]), int(rp[b])
        ref, _ = oracle_ref("fp64", b - a, 1000, rp[a:b + 1] - e0, col[e0:e1] - a, w[e0:e1], gcal[a:b], iapp[a:b],
                            stride=50)
        assert np.abs(out[:, a:b] - ref).max() < 5e-4, k


def test_nets_refuses_what_one_workgroup_cannot_hold():
    """A connected network larger than a workgroup holds (2,048 cells in
    FP32, 1,024 in FP64) is refused with INFOLI_ERR_ARG, and so is the
    gather in a launch mode other than persistent; the strategy in force
    stays."""
    for prec, n in (("fp32k", 2100), ("fp64", 1100)):
        rp, col, w = I.gen_symmetric(3, n, 0.02, 0.04)
        net = make_net(n, prec, rp, col, w)
        before = net.exec
        with pytest.raises(I.InfoliError) as ei:
            net.set_exec(**NETS)
        assert ei.value.code == I.ERR_ARG
        assert net.exec == before
        net.close()
    rp, col, w = I.gen_symmetric(3, 300, 0.05, 0.04)
    net = make_net(300, "fp32k", rp, col, w)
    with pytest.raises(I.InfoliError):
        net.set_exec(mode=I.EXEC_GRAPH, gather=I.GATHER_NETS)
    net.close()
