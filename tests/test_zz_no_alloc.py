"""The entry point that allocates nothing, infoli_gap_currents_noalloc,
against infoli_gap_currents through the C ABI: equal bit for bit on two
calls in a row on one GPU, the second with a smaller network (and another
precision), so the shared static scratch is reused with stale contents
beyond the second network's cells; a network beyond the scratch limit is
refused with INFOLI_ERR_ARG, and the refusal leaves the next call intact."""
import numpy as np
import pytest

import infoli_amd as I
from test_gpu_parity import make_net, say

pytestmark = pytest.mark.gpu


def _net(n, prec, p, seed):
    rp, col, w = I.gen_symmetric(seed, n, p, 0.02, 0.06)
    gcal = I.gen_uniform(seed, 2, n, 0.5, 1.1)
    iapp = I.gen_uniform(seed, 3, n, -1.0, 3.0)
    net = make_net(n, prec, rp, col, w, gcal, iapp)
    net.advance(40)   # voltages apart, so the currents are not zero
    return net


def test_gap_currents_noalloc_equals_allocating_entry_point():
    say("gap currents without allocation")
    calls = [(6000, "fp32k", 0.02, 5), (1500, "fp64", 0.05, 6), (700, "fp32", 0.1, 7)]
    for n, prec, p, seed in calls:
        net = _net(n, prec, p, seed)
        ref = net.gap_currents()
        out = net.gap_currents(noalloc=True)
        again = net.gap_currents(noalloc=True)
        net.close()
        say(f"  {n} cells {prec}: max |I_c| {np.abs(ref).max():.3e}")
        assert np.abs(ref).max() > 0
        assert np.array_equal(out, ref) and np.array_equal(again, ref)


def test_gap_currents_noalloc_refuses_beyond_its_scratch():
    say("gap currents without allocation: limit")
    n = 2 * 1024 * 1024 + 64           # FP64: 8 bytes per cell, more than 16 MiB
    big = I.Network(n, precision="fp64")
    with pytest.raises(I.InfoliError) as e:
        big.gap_currents(noalloc=True)
    assert e.value.code == -1
    big.close()
    net = _net(900, "fp64", 0.05, 8)
    assert np.array_equal(net.gap_currents(noalloc=True), net.gap_currents())
    net.close()
