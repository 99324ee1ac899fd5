"""Seeded synthetic workloads: the five benchmark configurations.

Every input comes from the library's counter-based generators keyed by
(master seed, stream, index), so the GPU run, the CPU oracle and any later
run see bit-identical inputs.  Streams: 0 graph, 1 w, 2 g_CaL, 3 Iapp,
4 pulse mask, 5 network sizes.

  plain_1k     1,000 cells, symmetric G(n, 0.1), w = 0.04, g_CaL 0.68,
               6 uA/cm^2 pulse to every cell on steps [10000, 20000),
               20,000 steps (the reference's own CPU-runnable case).
  dense_10k    10,000 cells, G(n, 0.1), w = 0.04, g_CaL ~ U[0.5, 1.1],
               Iapp ~ N(0, 0.5^2) constant, 10,000 steps.
  sparse_100k  100,000 cells, G(n, 100/(n-1)) (degree ~ Poisson(100)),
               w ~ U[0.02, 0.06], g_CaL ~ U[0.5, 1.1], 6 uA/cm^2 pulse to a
               random 20% of cells on steps [2000, 4000), 10,000 steps.
  batched_512  512 networks of round(exp(U[ln 64, ln 4096])) cells, each
               G(N_k, 0.1), w = 0.04, one g_CaL ~ U[0.5, 1.1] and one
               Iapp ~ U[-1, 3] per network, 20,000 steps.
  large_1m     1,000,000 cells, G(n, 50/(n-1)), w ~ U[0.02, 0.06],
               g_CaL ~ U[0.5, 1.1], Iapp ~ N(1, 1) constant, 2,000 steps.

And a batch of small networks, the shape a parameter sweep takes:

  sweep_256    a parameter sweep of small networks: 16 x 16 networks of
               256 cells, G(256, 0.1), w = 0.04, g_CaL on a 16-point grid
               of [0.5, 1.1] and constant Iapp on a 16-point grid of
               [-1, 3] uA/cm^2, 20,000 steps (the batching the reference's
               per-process runs would need 256 processes for).

And one network for the precision statement where the gap junctions act
on spikes (the five above either stay subthreshold or, in plain_1k and
batched_512, keep every cell of a network identical, so I_c = 0):

  spiking_20k  20,000 cells, G(n, 50/(n-1)), w ~ U[0.02, 0.06],
               g_CaL ~ U[0.9, 1.4], Iapp ~ U[-1.2, -0.6] constant, 6,000
               steps.  Isolated cells of these parameters fire
               persistently; coupled, they fire from about 100 ms on, at
               different times, so every spike drives current through its
               junctions.
"""
import math

import numpy as np

from .network import gen_mask, gen_normal, gen_symmetric, gen_uniform

MASTER_SEED = 20261203
S_GRAPH, S_W, S_GCAL, S_IAPP, S_MASK, S_SIZES = 0, 1, 2, 3, 4, 5

NAMES = ("plain_1k", "dense_10k", "sparse_100k", "batched_512", "large_1m", "spiking_20k", "sweep_256")


def _net_seed(seed, k):
    return (seed ^ ((k + 1) * 0x9E3779B97F4A7C15)) & 0xFFFFFFFFFFFFFFFF


class Workload(dict):
    """n, row_ptr, col, w, gcal, iapp_base, mask, amp, t0, t1, steps,
    partition (None or offsets) and a description."""

    @property
    def nnz(self):
        return int(self["row_ptr"][-1])


def plain_1k(n=1000, steps=20000, seed=MASTER_SEED):
    rp, col, w = gen_symmetric(seed, n, 0.1, 0.04)
    return Workload(name="plain_1k", n=n, row_ptr=rp, col=col, w=w,
                    gcal=np.full(n, 0.68), iapp_base=None, mask=None, amp=6.0,
                    t0=10000, t1=20000, steps=steps, partition=None)


def dense_10k(n=10000, steps=10000, seed=MASTER_SEED, p=0.1):
    rp, col, w = gen_symmetric(seed, n, p, 0.04)
    return Workload(name="dense_10k", n=n, row_ptr=rp, col=col, w=w,
                    gcal=gen_uniform(seed, S_GCAL, n, 0.5, 1.1),
                    iapp_base=gen_normal(seed, S_IAPP, n, 0.0, 0.5), mask=None, amp=0.0,
                    t0=0, t1=0, steps=steps, partition=None)


def sparse_100k(n=100000, steps=10000, seed=MASTER_SEED, mean_deg=100.0):
    rp, col, w = gen_symmetric(seed, n, mean_deg / (n - 1), 0.02, 0.06)
    return Workload(name="sparse_100k", n=n, row_ptr=rp, col=col, w=w,
                    gcal=gen_uniform(seed, S_GCAL, n, 0.5, 1.1), iapp_base=None,
                    mask=gen_mask(seed, S_MASK, n, 0.2), amp=6.0, t0=2000, t1=4000,
                    steps=steps, partition=None)


def batched_512(n_nets=512, lo=64, hi=4096, steps=20000, seed=MASTER_SEED, p=0.1):
    u = gen_uniform(seed, S_SIZES, n_nets, math.log(lo), math.log(hi))
    sizes = np.array([int(round(math.exp(x))) for x in u], np.int64)
    gnet = gen_uniform(seed, S_GCAL, n_nets, 0.5, 1.1)
    inet = gen_uniform(seed, S_IAPP, n_nets, -1.0, 3.0)
    offs = np.zeros(n_nets + 1, np.int64)
    offs[1:] = np.cumsum(sizes)
    n = int(offs[-1])
    rps, cols, ws = [np.zeros(1, np.int64)], [], []
    base_nnz = 0
    for k in range(n_nets):
        rp, col, w = gen_symmetric(_net_seed(seed, k), int(sizes[k]), p, 0.04)
        rps.append(rp[1:] + base_nnz)
        cols.append(col + offs[k])
        ws.append(w)
        base_nnz += int(rp[-1])
    row_ptr = np.concatenate(rps)
    col = np.concatenate(cols).astype(np.int32)
    w = np.concatenate(ws)
    gcal = np.repeat(gnet, sizes)
    iapp = np.repeat(inet, sizes)
    return Workload(name="batched_512", n=n, row_ptr=row_ptr, col=col, w=w, gcal=gcal,
                    iapp_base=iapp, mask=None, amp=0.0, t0=0, t1=0, steps=steps,
                    partition=offs, sizes=sizes)


def sweep_256(grid=16, size=256, steps=20000, seed=MASTER_SEED, p=0.1):
    """A parameter sweep of small networks: grid x grid networks of `size`
    cells, network (a, b) with g_CaL = 0.5 + 0.6 a / (grid - 1) and a
    constant Iapp = -1 + 4 b / (grid - 1) uA/cm^2, each G(size, p) with
    w = 0.04."""
    n_nets = grid * grid
    offs = np.arange(n_nets + 1, dtype=np.int64) * size
    rps, cols, ws = [np.zeros(1, np.int64)], [], []
    base_nnz = 0
    for k in range(n_nets):
        rp, col, w = gen_symmetric(_net_seed(seed, k), size, p, 0.04)
        rps.append(rp[1:] + base_nnz)
        cols.append(col + offs[k])
        ws.append(w)
        base_nnz += int(rp[-1])
    a = np.repeat(np.arange(grid), grid)
    b = np.tile(np.arange(grid), grid)
    gnet = 0.5 + 0.6 * a / (grid - 1)
    inet = -1.0 + 4.0 * b / (grid - 1)
    sizes = np.full(n_nets, size, np.int64)
    return Workload(name="sweep_256", n=int(offs[-1]), row_ptr=np.concatenate(rps),
                    col=np.concatenate(cols).astype(np.int32), w=np.concatenate(ws),
                    gcal=np.repeat(gnet, size), iapp_base=np.repeat(inet, size), mask=None, amp=0.0,
                    t0=0, t1=0, steps=steps, partition=offs, sizes=sizes)


def large_1m(n=1000000, steps=2000, seed=MASTER_SEED, mean_deg=50.0):
    rp, col, w = gen_symmetric(seed, n, mean_deg / (n - 1), 0.02, 0.06)
    return Workload(name="large_1m", n=n, row_ptr=rp, col=col, w=w,
                    gcal=gen_uniform(seed, S_GCAL, n, 0.5, 1.1),
                    iapp_base=gen_normal(seed, S_IAPP, n, 1.0, 1.0), mask=None, amp=0.0,
                    t0=0, t1=0, steps=steps, partition=None)


def spiking_20k(n=20000, steps=6000, seed=MASTER_SEED, mean_deg=50.0):
    rp, col, w = gen_symmetric(seed, n, mean_deg / (n - 1), 0.02, 0.06)
    return Workload(name="spiking_20k", n=n, row_ptr=rp, col=col, w=w,
                    gcal=gen_uniform(seed, S_GCAL, n, 0.9, 1.4),
                    iapp_base=gen_uniform(seed, S_IAPP, n, -1.2, -0.6), mask=None, amp=0.0,
                    t0=0, t1=0, steps=steps, partition=None)


BUILDERS = {"plain_1k": plain_1k, "dense_10k": dense_10k, "sparse_100k": sparse_100k,
            "batched_512": batched_512, "large_1m": large_1m, "spiking_20k": spiking_20k,
            "sweep_256": sweep_256}


def build(name, **kw):
    return BUILDERS[name](**kw)


def load_into(net, wl):
    """Copy a workload's inputs into a Network."""
    net.set_connectivity(wl["row_ptr"], wl["col"], wl["w"])
    if wl.get("partition") is not None:
        net.set_partition(wl["partition"])
    net.set_gcal(wl["gcal"])
    net.set_iapp(wl["iapp_base"])
    net.set_iapp_pulse(wl["amp"], wl["t0"], wl["t1"], wl["mask"])
