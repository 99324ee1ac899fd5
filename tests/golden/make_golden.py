#!/usr/bin/env python3
"""Regenerate the golden fixtures in this directory from the reference.

Runs the reference's own programs, compiled from its sources by
oracle/Makefile into oracle/_ref/ (needs the reference mounted at
/root/reference; the fixtures are committed so nothing else does):

  st.x      src/single_threaded/infoli.c   FP64 simulator.  Given an Iapp file
            it runs one step per line and prints V_axon of every cell with
            %.16f (ST:242-286); connectivity comes from cellConnections.txt,
            a dense [sender][receiver] matrix (ST:158-191); g_CaL = 0.68.
  stable.x  src/regression_test/stable_version.c with -DG_CAL_FROM_FILE:
            the regression baseline writer (float state, double exp/pow),
            10 cells, 40,000 steps, %.8f per cell per step.

Also copies the data files the reference's regression test uses
(bin/default.conf, src/regression_test/cellConnections.txt, gcal_file.txt).

Usage: python tests/golden/make_golden.py   (from the repository root)
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
REFBIN = os.path.join(ROOT, "oracle", "_ref")


def run_st(n, conn, iapp):
    """Run st.x on an n x n [sender][receiver] matrix and an Iapp table
    [steps][n]; return the FP64 V_axon trace [steps][n]."""
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "cellConnections.txt"), "w") as fh:
            for row in conn:
                fh.write(" ".join("%.4f" % v for v in row) + "\n")
        with open(os.path.join(d, "iapp.txt"), "w") as fh:
            for row in iapp:
                fh.write(" ".join("%.3f" % v for v in row) + "\n")
        subprocess.run([os.path.join(REFBIN, "st.x"), str(n), "iapp.txt"], cwd=d,
                       check=True, stdout=subprocess.DEVNULL)
        lines = open(os.path.join(d, "InferiorOlive_Output.txt")).read().splitlines()
    rows = []
    for ln in lines[1:]:
        tok = ln.split()
        rows.append([float(x) for x in tok[2:2 + n]])
    tr = np.array(rows, np.float64)
    assert tr.shape == (len(iapp), n), tr.shape
    return tr


def conn_as_read(conn):
    """The doubles the reference reads back from the %.4f text."""
    return np.array([[float("%.4f" % v) for v in row] for row in conn], np.float64)


def iapp_as_read(iapp):
    return np.array([[float("%.3f" % v) for v in row] for row in iapp], np.float64)


def main():
    for f in ("st.x", "stable.x"):
        if not os.path.exists(os.path.join(REFBIN, f)):
            sys.exit(f"missing {REFBIN}/{f}: run make -C oracle ref")
    shutil.copyfile(os.path.join(REF, "bin", "default.conf"), os.path.join(HERE, "default.conf"))
    for f in ("cellConnections.txt", "gcal_file.txt"):
        shutil.copyfile(os.path.join(REF, "src", "regression_test", f), os.path.join(HERE, f))

    rng = np.random.default_rng(20261203)

    # 1. single cell, 4000 steps (200 ms): rest, then a 25 ms 6 uA/cm^2 pulse
    #    (drives a spike), then recovery.  Pins the compartment update alone.
    n = 1
    iapp = np.zeros((4000, n))
    iapp[1000:1500, :] = 6.0
    conn = np.zeros((n, n))
    tr = run_st(n, conn, iapp)
    np.savez_compressed(os.path.join(HERE, "st_single_cell.npz"),
                        conn=conn_as_read(conn), iapp=iapp_as_read(iapp), trace=tr)

    # 2. the regression set-up's 10-cell complete graph (w = 0.04), cells made
    #    distinct by a constant per-cell Iapp 0, 0.5, ..., 4.5 (with equal
    #    drive the cells stay synchronous and I_c is identically 0).
    n = 10
    conn = np.loadtxt(os.path.join(HERE, "cellConnections.txt"))
    iapp = np.tile(0.5 * np.arange(n), (2000, 1))
    tr = run_st(n, conn, iapp)
    np.savez_compressed(os.path.join(HERE, "st_reg10_iapp.npz"),
                        conn=conn_as_read(conn), iapp=iapp_as_read(iapp), trace=tr)

    # 3. 20 cells, directed random graph (density 0.3), w in [0.01, 0.08],
    #    per-cell Iapp in [-1, 3] plus a 6 uA/cm^2 pulse to 5 cells on steps
    #    [300, 600).  1500 steps (75 ms).
    n = 20
    conn = np.where(rng.random((n, n)) < 0.3, rng.uniform(0.01, 0.08, (n, n)), 0.0)
    np.fill_diagonal(conn, 0.0)
    base = rng.uniform(-1.0, 3.0, n)
    iapp = np.tile(base, (1500, 1))
    pulsed = rng.choice(n, 5, replace=False)
    iapp[300:600, pulsed] += 6.0
    tr = run_st(n, conn, iapp)
    np.savez_compressed(os.path.join(HERE, "st_rand20.npz"),
                        conn=conn_as_read(conn), iapp=iapp_as_read(iapp), trace=tr)

    # 4. the reference regression baseline: stable.x 10 (40,000 steps, the
    #    hard-coded pulse on steps [20000, 20499)); the first 10,000 steps
    #    are the "standard" strength of regression_test.sh.  Kept every 4th
    #    step as float32 (the %.8f text round-trips float32).
    with tempfile.TemporaryDirectory() as d:
        for f in ("cellConnections.txt", "gcal_file.txt"):
            shutil.copy(os.path.join(HERE, f), d)
        subprocess.run([os.path.join(REFBIN, "stable.x"), "10"], cwd=d, check=True,
                       stdout=subprocess.DEVNULL)
        base_tr = np.loadtxt(os.path.join(d, "InferiorOlive_Output_Baseline.txt"),
                             dtype=np.float64)
    assert base_tr.shape == (40000, 10), base_tr.shape
    np.savez_compressed(os.path.join(HERE, "stable_reg10.npz"),
                        step_index=np.arange(0, 10000, 4),
                        trace=base_tr[0:10000:4].astype(np.float32))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
