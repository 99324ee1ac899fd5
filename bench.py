#!/usr/bin/env python3
"""Benchmark: cell-steps/s of the InfOli network update on one MI355X.

Default workload: sparse_100k (100,000 cells, degree ~ Poisson(100),
w ~ U[0.02, 0.06], g_CaL ~ U[0.5, 1.1], 6 uA/cm^2 pulse to 20% of cells on
steps [2000, 4000)), FP32 arithmetic with compensated gap-junction sums,
V_axon trace every 100 steps (the reference's print granularity).  A step is
one timestep of the whole network; --steps K of them are timed after
--warmup W untimed ones, inputs resident in HBM.

Printed (one JSON line): the metric, the roofline of the step kernel
(algorithmic bytes per launch / mean launch time from HIP events on the
library's stream) and the CPU baseline (the oracle's OpenMP FP32
restatement of the reference loops, timed on this host's cores on a bounded
sample of the same workload).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# hardware exp2 (v_exp_f32) throughput measured on the MI355X by
# `infoli_amd/bin/infoli_calib.x exp`: the ceiling of the gap-junction terms,
# one exp each, when the sums run from on-chip memory (the dense network)
VEXP_PEAK_PER_S = 13.0e12
GATHER_NAMES = {0: "direct", 1: "lds", 2: "staged", 3: "tiled", 4: "sweep", 5: "nets"}
TILED_KERNELS = {0: "k_tile_gj+k_tile_ode", 1: "k_tile_gj+k_tile_ode3", 2: "k_tile_gj+k_tile_dend"}
SWEEP_KERNELS = {0: "k_tile_sweep", 1: "k_tile_sweep+k_tile_ode3", 2: "k_tile_sweep+k_tile_beside+k_tile_dend"}


def kernel_name(exec_cfg):
    """The kernels of one step under this execution strategy."""
    if exec_cfg["gather"] == 3:
        return TILED_KERNELS[exec_cfg["ode_split"]]
    if exec_cfg["gather"] == 4:
        return SWEEP_KERNELS[exec_cfg["ode_split"]]
    if exec_cfg["gather"] == 5:
        return "k_persist_nets"
    return "k_persist" if exec_cfg["mode"] == 3 else "k_step"


def measured_traffic(kernel, workload, prec):
    """HBM bytes per step from the committed rocprofv3 --pmc summary of this
    kernel and workload (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel
    of known size), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_{prec}.md")
    try:
        for line in open(path):
            if line.startswith("| per step |") and "total" in line:
                return float(line.split("total")[1].split("MB")[0]) * 1e6
    except OSError:
        pass
    return None
METRIC = "cell-steps/s at 100k cells, FP32+Kahan batched vs FP64 plain; % roofline"


def algorithmic_bytes_per_step(n, nnz, prec):
    """Bytes one step must move: per neighbour entry the sender id (int32)
    and conductance (4 or 8 B); per cell 16 state values read and written,
    g_CaL and Iapp read, V_dend read once, V_dend / trace written."""
    t = 8 if prec == "fp64" else 4
    per_entry = 4 + t
    per_cell = 36 * t
    return per_entry * nnz + per_cell * n


def _bind_openmp_threads():
    """One OpenMP thread per core, packed (OMP_PLACES=cores,
    OMP_PROC_BIND=close), unless the caller chose otherwise.  libgomp reads
    these when it loads, so they only take effect before the first OpenMP
    library is loaded into this process."""
    try:
        if "libgomp" in open("/proc/self/maps").read():
            return False
    except OSError:
        return False
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    return True


def cpu_baseline(wl, target_s=18.0):
    """The oracle's OpenMP restatement of the reference step on this host's
    cores: the FP32 rate is the median of three runs over steps [0, k) with k
    sized to about target_s / 4 seconds each; one FP64 run of the same steps
    is timed beside it.  Returns (cpu_baseline object, FP64 rate)."""
    bound = _bind_openmp_threads()
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    n = wl["n"]
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, int(os.environ.get("OMP_NUM_THREADS", threads))))

    def run(prec, k):
        st = O.initial_state(n)
        st[O.G_CAL] = wl["gcal"]
        _, el = O.run(prec, st, k, wl["row_ptr"], wl["col"], wl["w"],
                      iapp_base=wl["iapp_base"], mask=wl["mask"], amp=wl["amp"],
                      t0=wl["t0"], t1=wl["t1"], stride=100, nthreads=threads, trace=True)
        return el

    probe = run(O.FP32, 2)
    k = int(max(2, min(2000, target_s / 4 / max(probe / 2, 1e-6))))
    runs = sorted(run(O.FP32, k) for _ in range(3))
    el = runs[1]
    el64 = run(O.FP64, k)
    cpu = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    binding = "OMP_PLACES=cores OMP_PROC_BIND=close" if bound else "OpenMP binding as inherited"
    return ({"value": n * k / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
             "sample": f"{wl['name']}: steps [0, {k}) of {wl['steps']}, all {n} cells, FP32 "
                       f"OpenMP restatement of src/openmp/infoli.c:496-679 (oracle/infoli_oracle.c, "
                       f"gcc -O3 -march=x86-64-v3), {threads} OpenMP threads = the logical CPUs in this "
                       f"process's affinity mask (not the host's physical core count), on {cpu}, "
                       f"{binding}; median of 3 runs: {', '.join(f'{r:.2f}' for r in runs)} s"},
            n * k / el64)


def graph_steps(steps, default):
    """Steps per graph replay for a timed run of `steps`: the largest divisor
    of it not above `default`, so the timed steps are whole replays of one
    graph (captured before the timed region)."""
    return max(d for d in range(1, min(steps, default) + 1) if steps % d == 0)


def layout_bytes_per_step(n, nnz, prec, constant_w):
    """Bytes one step moves in the tiled layout itself: per neighbour entry a
    16-bit chunk-local sender id, plus the conductance unless the tile's
    conductance is constant (then none); per cell the 144 (FP32) or 288 B of
    state, stimulus and V_dend of the algorithmic count."""
    t = 8 if prec == "fp64" else 4
    return (2 + (0 if constant_w else t)) * nnz + 36 * t * n


def batched_leg(I, W, steps, warmup, per_net_steps, per_net_warmup, with_fp64=True):
    """batched_512 in FP32 with compensated sums: all 512 networks in one
    context (block-diagonal), 500-step graph replays; beside it the same
    networks one after another, each in its own FP64 context."""
    wl = W.batched_512()
    n, nnz = wl["n"], wl.nnz
    net = I.Network(n, precision="fp32k")
    W.load_into(net, wl)
    G = graph_steps(steps, 500)
    net.set_exec(steps_per_launch=G)
    net.prepare(steps, 100)
    net.advance(warmup, trace_stride=100)
    net.synchronize()
    t0 = time.perf_counter()
    net.advance(steps, trace_stride=100)
    net.synchronize()
    el = time.perf_counter() - t0
    dev_ms = net.last_timing()["advance_ms"]
    launches = net.last_launches()
    exec_cfg = net.exec
    net.close()
    value = n * steps / el
    step_ms = dev_ms / steps
    lb = layout_bytes_per_step(n, nnz, "fp32k", constant_w=True)
    out = {"workload": "batched_512", "networks": len(wl["partition"]) - 1, "cells": n,
           "neighbour_entries": nnz, "steps": steps, "warmup": warmup,
           "cell_steps_per_s": value, "ms_per_step": el * 1e3 / steps, "exec": exec_cfg,
           "launch_path": dict(launches, steps_per_graph=G),
           "roofline": {"bound": "hbm", "bytes_per_step": lb, "achieved": lb / (step_ms * 1e-3) / 1e9,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": lb / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "note": "bytes of the tiled layout itself: 2 B per entry (16-bit ids; the "
                                "conductance 0.04 is constant per tile and not stored) + 144 B per cell"}}
    if with_fp64:
        print("# batched_512: per-network FP64 launches", file=sys.stderr, flush=True)
        v, desc = per_network_fp64(I, wl, per_net_steps, per_net_warmup)
        out["per_network_fp64_cell_steps_per_s"] = v
        out["per_network_fp64"] = desc
        out["batched_over_per_network_fp64"] = value / v
    return out


def per_network_fp64(I, wl, steps, warmup):
    """The batched workload's networks run one after another, each in its own
    FP64 context (plain sums), as a per-network launch would: returns
    cell-steps/s over the summed per-network times and a description."""
    offs, rp, col, w = wl["partition"], wl["row_ptr"], wl["col"], wl["w"]
    total_s, cells = 0.0, 0
    for k in range(len(offs) - 1):
        a, b = int(offs[k]), int(offs[k + 1])
        e0, e1 = int(rp[a]), int(rp[b])
        net = I.Network(b - a, precision="fp64")
        net.set_connectivity(rp[a:b + 1] - e0, col[e0:e1] - a, w[e0:e1])
        net.set_gcal(wl["gcal"][a:b])
        net.set_iapp(wl["iapp_base"][a:b])
        net.set_exec(steps_per_launch=graph_steps(steps, 500))
        net.prepare(steps)
        net.advance(warmup)
        net.synchronize()
        t0 = time.perf_counter()
        net.advance(steps)
        net.synchronize()
        total_s += time.perf_counter() - t0
        cells += b - a
        net.close()
    return cells * steps / total_s, (f"{len(offs) - 1} networks, each its own FP64 context "
                                     f"(heuristic launch strategy), {steps} timed steps after "
                                     f"{warmup} warm-up, run one after another; {total_s:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--workload", default="sparse_100k")
    ap.add_argument("--precision", default="fp32k", choices=["fp32k", "fp32", "fp64"])
    ap.add_argument("--stride", type=int, default=100)
    ap.add_argument("--mode", default="auto", choices=["auto", "graph", "persistent", "step", "tune"])
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--gather", default="auto", choices=["auto", "direct", "lds", "staged", "tiled", "sweep", "nets"])
    ap.add_argument("--ode-split", type=int, default=-1,
                    help="compartment placement with the tiled gathers (0, 1, 2; default: heuristic)")
    ap.add_argument("--lds-kb", type=int, default=0)
    ap.add_argument("--tile-cells", type=int, default=0)
    ap.add_argument("--steps-per-launch", type=int, default=0,
                    help="steps per graph launch (default 100; 500 for batched_512)")
    ap.add_argument("--study", action="store_true",
                    help="precision study instead of the benchmark: FP32 compensated and FP32 "
                         "plain against FP64, max |dV| over 100 ms and spike times after it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched_512 leg (all networks in one context vs per-network FP64)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              f"(sharding a network over GPUs by cell ranges is designed, not implemented)", file=sys.stderr)
        sys.exit(2)

    import infoli_amd as I
    from infoli_amd import workloads as W

    print(f"# building workload {args.workload}", file=sys.stderr, flush=True)
    wl = W.build(args.workload)
    n, nnz = wl["n"], wl.nnz

    def make(prec):
        net = I.Network(n, precision=prec)
        W.load_into(net, wl)
        if args.lanes:
            net.set_exec(lanes_per_cell=args.lanes)
        if args.lds_kb:
            net.set_exec(lds_kb=args.lds_kb)
        if args.tile_cells:
            net.set_exec(tile_cells=args.tile_cells)
        spl = args.steps_per_launch or graph_steps(args.steps, 500 if wl.get("partition") is not None else 100)
        net.set_exec(steps_per_launch=spl)
        if args.gather != "auto" or args.ode_split >= 0:
            g = {"direct": 0, "lds": 1, "staged": 2, "tiled": 3, "sweep": 4, "nets": 5}.get(args.gather)
            net.set_exec(gather=g, ode_split=args.ode_split if args.ode_split >= 0 else None,
                         mode=I.EXEC_PERSISTENT if g == 5 else None)
        if args.mode == "graph":
            net.set_exec(mode=I.EXEC_GRAPH)
        elif args.mode == "persistent":
            net.set_exec(mode=I.EXEC_PERSISTENT, steps_per_launch=500)
        elif args.mode == "step":
            net.set_exec(mode=I.EXEC_STEP)
        elif args.mode == "tune":
            net.autotune(trial_steps=200)
        return net

    def timed(net, steps, warmup):
        # the graph the timed steps replay is captured here, before the
        # warm-up and the timed region
        net.prepare(steps, args.stride)
        if warmup:
            net.advance(warmup, trace_stride=args.stride)
        net.synchronize()
        t0 = time.perf_counter()
        net.advance(steps, trace_stride=args.stride)
        net.synchronize()
        return time.perf_counter() - t0

    if args.study:
        from infoli_amd import analysis
        steps = args.steps if args.steps != ap.get_default("steps") else wl["steps"]
        res = {"study": "precision", "workload": wl["name"], "cells": n, "neighbour_entries": nnz,
               "reference": "fp64"}
        for prec in ("fp32k", "fp32"):
            print(f"# study {prec} vs fp64, {steps} steps", file=sys.stderr, flush=True)
            a, b = make(prec), make("fp64")
            res[prec] = analysis.compare_runs(
                a, b, steps, progress=lambda t, dv: print(f"#   step {t}: max |dV| so far {dv:.3e} mV",
                                                          file=sys.stderr, flush=True))
            res[prec]["exec"] = a.exec
            a.close()
            b.close()
        print(json.dumps(res), flush=True)
        return

    print(f"# {args.precision}: {n} cells, {nnz} neighbour entries", file=sys.stderr, flush=True)
    net = make(args.precision)
    exec_cfg = net.exec
    el = timed(net, args.steps, args.warmup)
    value = n * args.steps / el
    dev_ms = net.last_timing()["advance_ms"]
    launches = dict(net.last_launches(), steps_per_graph=net.exec["steps_per_launch"])

    # Kernel timing: HIP events recorded on the library's stream around the
    # timed region (infoli_last_timing); one step's launches (the tiled
    # gather's two kernels, or the fused step kernel) are the unit.  A second
    # figure brackets every step separately (launch mode STEP, so the
    # launches are not overlapped by the graph) for comparison.
    mean_launch_ms = dev_ms / args.steps
    steps_per_launch = 1.0
    bytes_per_launch = algorithmic_bytes_per_step(n, nnz, args.precision)
    bytes_note = "8 B per neighbour entry (int32 id + conductance) + 144 B per cell"
    if exec_cfg["gather"] in (3, 4) and nnz > 0 and bool(np.all(wl["w"] == wl["w"][0])):
        # tiles of one conductance everywhere (the batched and dense
        # networks, w = 0.04) hold 16-bit ids and no conductances: the step
        # is priced against its own layout, not the 8 B count it beats
        bytes_per_launch = layout_bytes_per_step(n, nnz, args.precision, constant_w=True)
        bytes_note = "tiled layout's own bytes: 2 B per entry (16-bit ids, constant conductance) + 144 B per cell"
    achieved = bytes_per_launch / (mean_launch_ms * 1e-3) / 1e9
    net.set_profiling(True)
    prof_steps = min(args.steps, 2000)
    net.advance(prof_steps)
    tm = net.last_timing()
    net.set_profiling(False)
    per_step_bracketed_ms = tm["kernel_ms"] / max(tm["kernel_launches"], 1)
    net.close()

    kname = kernel_name(exec_cfg)
    tr = measured_traffic(kname, wl["name"], args.precision)
    traffic = None if tr is None else tr * steps_per_launch
    exp_rate = nnz / (mean_launch_ms * 1e-3)
    extra = {"device_ms": round(dev_ms, 3), "exec": exec_cfg, "launch_path": launches,
             "mean_step_launches_ms": round(mean_launch_ms, 5),
             "per_step_bracketed_ms": round(per_step_bracketed_ms, 5),
             "exp_rate": {"achieved": exp_rate, "peak": VEXP_PEAK_PER_S, "unit": "exp/s",
                          "frac": exp_rate / VEXP_PEAK_PER_S,
                          "note": "one exp per neighbour entry per step; peak = measured v_exp_f32 rate"}}
    if not args.no_fp64 and args.precision != "fp64":
        print("# fp64 comparison run", file=sys.stderr, flush=True)
        n64 = make("fp64")
        k64 = min(args.steps, 2000)
        el64 = timed(n64, k64, min(args.warmup, 200))
        n64.close()
        extra["fp64_cell_steps_per_s"] = n * k64 / el64
        extra["speedup_vs_fp64"] = value / extra["fp64_cell_steps_per_s"]

    if wl.get("partition") is not None and not args.no_fp64:
        print("# per-network FP64 launches", file=sys.stderr, flush=True)
        v, desc = per_network_fp64(I, wl, min(args.steps, 500), min(args.warmup, 100))
        extra["per_network_fp64_cell_steps_per_s"] = v
        extra["per_network_fp64"] = desc
        extra["batched_over_per_network_fp64"] = value / v
    elif not args.no_batched:
        # a bounded batched leg beside the headline: 1,000 steps (two 500-step
        # graph replays) after 500, against 500 per-network FP64 steps each
        print("# batched_512 leg", file=sys.stderr, flush=True)
        extra["batched"] = batched_leg(I, W, 1000, 500, 500, 100, with_fp64=not args.no_fp64)

    cpu = None
    if not args.no_cpu:
        print("# CPU baseline", file=sys.stderr, flush=True)
        cpu, cpu64 = cpu_baseline(wl)
        extra["gpu_over_cpu"] = value / cpu["value"]
        extra["cpu_fp64_cell_steps_per_s"] = cpu64

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded counter-based generators)",
        "config": {"workload": wl["name"], "cells": n, "neighbour_entries": nnz,
                   "mean_degree": round(nnz / n, 3), "precision": {
                       "fp32k": "fp32 + compensated gap-junction sums", "fp32": "fp32",
                       "fp64": "fp64"}[args.precision],
                   "trace_stride": args.stride, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_launch": bytes_per_launch, "bytes_counted": bytes_note,
                     "launch": "one step (its kernels, back to back in the graph)",
                     "kernel": kname},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
