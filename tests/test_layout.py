"""Host-only checks of the 2D-tiled layout (no GPU): the tables the library
builds for a network against the network itself, and one step's scratch
traffic as the kernels index it (every partial-sum and dendrite-sum slot a
compartment kernel reads is written by the launch before it; every soma and
axon updated once).  The shapes include the one of the swept-tile test that
stopped with a NaN once (6,000 cells, 512-cell blocks, 2 KiB chunks, one
lane, compartments in the sweeping kernel)."""
import numpy as np
import pytest

import infoli_amd as I


def isolated_blocks_case(n=6000, p=0.01, seed=41, lo=1024, hi=2048):
    """The swept-tile test's network: a random symmetric graph whose cells
    [lo, hi) are isolated, so two 512-cell receiver blocks have no tiles."""
    rp, col, w = I.gen_symmetric(seed, n, p, 0.02, 0.06)
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ~(((rows >= lo) & (rows < hi)) | ((col >= lo) & (col < hi)))
    rp2 = np.zeros(n + 1, np.int64)
    rp2[1:] = np.cumsum(np.bincount(rows[keep], minlength=n))
    return rp2, col[keep].astype(np.int32), w[keep]


def block_diagonal_case(sizes=(70, 300, 1, 130, 900, 64), p=0.2, seed=5):
    rps, cols, ws, base, off = [np.zeros(1, np.int64)], [], [], 0, 0
    for k, m in enumerate(sizes):
        rp, col, w = I.gen_symmetric(seed + k, m, p, 0.04)
        rps.append(rp[1:] + base)
        cols.append(col + off)
        ws.append(w)
        base += int(rp[-1])
        off += m
    return np.concatenate(rps), np.concatenate(cols).astype(np.int32), np.concatenate(ws)


GATHER_SPLITS = [(I.GATHER_TILED, 0), (I.GATHER_TILED, 1), (I.GATHER_TILED, 2),
                 (I.GATHER_SWEEP, 0), (I.GATHER_SWEEP, 1), (I.GATHER_SWEEP, 2)]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32k"])
@pytest.mark.parametrize("L", [1, 4])
@pytest.mark.parametrize("R,kb", [(512, 2), (8192, 4)])
@pytest.mark.parametrize("gather,split", GATHER_SPLITS)
def test_swept_test_shapes(prec, L, R, kb, gather, split):
    rp, col, w = isolated_blocks_case()
    I.tiled_layout_check(rp, col, w, prec, gather, L, R, kb, split, 1024, nsa=7)


@pytest.mark.parametrize("nsa", [1, 2, 5, 12, 1000])
@pytest.mark.parametrize("block", [256, 512, 1024])
def test_soma_axon_workgroups_cover_every_cell(nsa, block):
    rp, col, w = isolated_blocks_case()
    I.tiled_layout_check(rp, col, w, "fp32k", I.GATHER_TILED, 1, 512, 2, 2, block, nsa)


@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
@pytest.mark.parametrize("gather,split", GATHER_SPLITS)
def test_block_diagonal_compact_tiles(prec, gather, split):
    """Blocks whose senders fit one chunk stage just their own range."""
    rp, col, w = block_diagonal_case()
    for R, kb in ((64, 48), (256, 2), (512, 48)):
        I.tiled_layout_check(rp, col, w, prec, gather, 2, R, kb, split, 1024, nsa=3)


@pytest.mark.parametrize("gather,split", GATHER_SPLITS)
def test_degenerate_networks(gather, split):
    for n, p in ((1, 0.0), (37, 0.0), (100, 1.0), (1100, 0.002)):
        rp, col, w = I.gen_symmetric(3, n, p, 0.02, 0.06)
        I.tiled_layout_check(rp, col, w, "fp32k", gather, 1, 64, 1, split, 1024, nsa=2)


def test_star_long_rows():
    """One hub receiving from 1,099 cells: long rows, wide slices."""
    n = 1100
    hub_row = np.arange(1, n, dtype=np.int32)
    rp = np.zeros(n + 1, np.int64)
    rp[1] = n - 1
    rp[2:] = n - 1 + np.arange(1, n)
    col = np.concatenate([hub_row, np.zeros(n - 1, np.int32)])
    w = np.full(len(col), 0.04)
    for L in (1, 2, 16):
        for gather, split in GATHER_SPLITS:
            I.tiled_layout_check(rp, col, w, "fp32", gather, L, 128, 1, split, 1024, nsa=4)


def test_check_rejects_bad_shapes():
    rp, col, w = isolated_blocks_case(n=600)
    for kw in (dict(lanes_per_cell=3), dict(tile_cells=100), dict(lds_kb=0), dict(ode_split=3),
               dict(gather=I.GATHER_SWEEP, block=512),
               dict(ode_split=2, nsa=0), dict(gather=I.GATHER_DIRECT)):
        with pytest.raises(I.InfoliError):
            I.tiled_layout_check(rp, col, w, **kw)
    bad = col.copy()
    bad[0] = 600
    with pytest.raises(I.InfoliError):
        I.tiled_layout_check(rp, bad, w)
