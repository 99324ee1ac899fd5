"""The CPU oracle against the reference's own outputs and the mathematics.

Pins (each a check a plausible restatement mistake would fail):
  - FP64 arithmetic equals the reference FP64 simulator (src/single_threaded,
    compiled from source into oracle/_ref/st.x) to the last printed digit on
    the committed golden traces;
  - FP32 step 1 from the default state equals the reference OpenMP
    simulator's printed value;
  - FP32 and FP64 both reproduce the reference regression baseline
    (src/regression_test/stable_version.c) within the reference checker's 1%
    criterion (regression_checker.c:44-66) over the 'standard' 500 ms;
  - gap-junction sums equal a long-double brute force, are antisymmetric and
    conserve sum_i I_c = 0 for symmetric conductances.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as O

CASES = ["st_single_cell", "st_reg10_iapp", "st_rand20"]


@pytest.mark.parametrize("case", CASES)
def test_fp64_equals_reference_st(golden, case):
    z = golden(case)
    conn, iapp, ref = z["conn"], z["iapp"], z["trace"]
    rp, col, w = O.dense_to_csr(conn)
    st = O.initial_state(conn.shape[0])
    out, _ = O.run(O.FP64, st, len(iapp), rp, col, w, table=iapp)
    fmt = np.vectorize(lambda x: "%.16f" % x)
    assert (fmt(out) == fmt(ref)).all()


def test_fp64_step1_known_answer():
    # ST prints -59.97596717 after the first step from the default state
    st = O.initial_state(1)
    out, _ = O.run(O.FP64, st, 1)
    assert "%.14f" % out[0, 0] == "-59.97596717300331"


def test_fp32_step1_known_answer():
    # the reference OpenMP simulator (float state) prints -59.97596741 after
    # step 1 from the default state with no input (every cell, any graph)
    st = O.initial_state(3)
    out, _ = O.run(O.FP32, st, 1, *O.dense_to_csr(np.full((3, 3), 0.04) - 0.04 * np.eye(3)))
    assert ["%.8f" % v for v in out[0]] == ["-59.97596741"] * 3


@pytest.mark.parametrize("case,tol", [("st_single_cell", 2e-3), ("st_reg10_iapp", 2e-4),
                                      ("st_rand20", 2e-2)])
def test_fp32_tracks_fp64_reference(golden, case, tol):
    z = golden(case)
    conn, iapp, ref = z["conn"], z["iapp"], z["trace"]
    rp, col, w = O.dense_to_csr(conn)
    for prec in (O.FP32, O.FP32K):
        st = O.initial_state(conn.shape[0])
        out, _ = O.run(prec, st, len(iapp), rp, col, w, table=iapp)
        err = np.abs(out - ref).max()
        assert err < tol, (prec, err)


def _regression_setup():
    conn = np.loadtxt(os.path.join(O.GOLDEN, "cellConnections.txt"))
    gcal = np.loadtxt(os.path.join(O.GOLDEN, "gcal_file.txt"))
    return O.dense_to_csr(conn, threshold=1e-7), gcal


def _checker_max_relative(base, test):
    # regression_checker.c:42-56: max over lines and cells of |d / base|
    b = base.astype(np.float32)
    t = test.astype(np.float32)
    return float(np.max(np.abs((b - t) / b)))


@pytest.mark.parametrize("prec", [O.FP64, O.FP32, O.FP32K])
def test_regression_baseline_standard(golden, prec):
    z = golden("stable_reg10")
    idx, base = z["step_index"], z["trace"]
    (rp, col, w), gcal = _regression_setup()
    st = O.initial_state(10)
    st[O.G_CAL] = gcal
    out, _ = O.run(prec, st, 10000, rp, col, w, amp=6.0, t0=20000, t1=20499)
    rel = _checker_max_relative(base, out[idx])
    assert rel <= 0.01, rel
    assert np.abs(out[idx] - base).max() < 5e-3


def test_gap_currents_bruteforce():
    rng = np.random.default_rng(3)
    n = 40
    conn = np.where(rng.random((n, n)) < 0.5, rng.uniform(0.01, 0.08, (n, n)), 0.0)
    np.fill_diagonal(conn, 0.0)
    rp, col, w = O.dense_to_csr(conn)
    v = rng.uniform(-70, 10, n)
    exact = np.zeros(n, np.longdouble)
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            d = np.longdouble(v[i]) - np.longdouble(v[col[e]])
            f = np.longdouble(0.8) * np.exp(-d * d / 100) + np.longdouble(0.2)
            exact[i] += np.longdouble(w[e]) * f * d
    ic64 = O.gap_currents(O.FP64, v, rp, col, w)
    scale = float(np.max(np.abs(exact))) + 1e-300
    assert np.max(np.abs(ic64 - exact.astype(np.float64))) <= 1e-13 * scale
    vf = v.astype(np.float32).astype(np.float64)
    exactf = np.zeros(n, np.longdouble)
    wf = w.astype(np.float32).astype(np.float64)
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            d = np.longdouble(vf[i]) - np.longdouble(vf[col[e]])
            exactf[i] += np.longdouble(wf[e]) * (np.longdouble(0.8) * np.exp(-d * d / 100) + 0.2) * d
    for prec in (O.FP32, O.FP32K):
        ic = O.gap_currents(prec, v, rp, col, w)
        assert np.max(np.abs(ic - exactf.astype(np.float64))) <= 2e-6 * scale, prec


def test_gap_currents_antisymmetric_and_conserved():
    rng = np.random.default_rng(5)
    n = 200
    a = np.triu(np.where(rng.random((n, n)) < 0.1, rng.uniform(0.02, 0.06, (n, n)), 0.0), 1)
    conn = a + a.T
    rp, col, w = O.dense_to_csr(conn)
    v = rng.uniform(-70, 0, n)
    ic = O.gap_currents(O.FP64, v, rp, col, w)
    total_abs = np.sum(np.abs(ic))
    assert abs(ic.sum()) <= 1e-13 * total_abs
    # odd in V: reflecting every voltage about any point negates I_c
    ic_r = O.gap_currents(O.FP64, -v, rp, col, w)
    assert np.allclose(ic_r, -ic, rtol=0, atol=1e-13 * np.max(np.abs(ic)))
    # equal voltages: no current, exactly
    assert (O.gap_currents(O.FP64, np.full(n, -60.0), rp, col, w) == 0).all()


def test_fp64_degree_budget_is_fp64_term_sum():
    """The degree budget of FP32K: at or above it, I_c is ST's FP64 sum over
    the FP32 values rounded once to float -- for float-exact voltages and
    conductances exactly (float)gj64, which the golden ST traces pin; a
    long-double brute force agrees to 1e-13.  Below it, the compensated FP32
    sum, unchanged."""
    rng = np.random.default_rng(11)
    n = 120
    conn = np.where(rng.random((n, n)) < rng.uniform(0.1, 0.9, (n, 1)), rng.uniform(0.02, 0.06, (n, n)), 0.0)
    np.fill_diagonal(conn, 0.0)
    rp, col, w = O.dense_to_csr(conn)
    deg = np.diff(rp)
    budget = int(np.median(deg))
    vf = rng.uniform(-70, 10, n).astype(np.float32).astype(np.float64)
    wf = w.astype(np.float32).astype(np.float64)
    try:
        O.set_fp64_degree(budget)
        hyb = O.gap_currents(O.FP32K, vf, rp, col, wf)
    finally:
        O.set_fp64_degree(0)
    plain = O.gap_currents(O.FP32K, vf, rp, col, wf)
    ic64 = O.gap_currents(O.FP64, vf, rp, col, wf)
    wide = deg >= budget
    assert 0 < wide.sum() < n
    assert np.array_equal(hyb[wide], ic64[wide].astype(np.float32).astype(np.float64))
    assert np.array_equal(hyb[~wide], plain[~wide])
    exact = np.zeros(n, np.longdouble)
    for i in np.nonzero(wide)[0]:
        for e in range(rp[i], rp[i + 1]):
            d = np.longdouble(vf[i]) - np.longdouble(vf[col[e]])
            exact[i] += np.longdouble(wf[e]) * (np.longdouble(0.8) * np.exp(-d * d / 100) + 0.2) * d
    scale = float(np.max(np.abs(exact))) + 1e-300
    assert np.max(np.abs(ic64[wide] - exact[wide].astype(np.float64))) <= 1e-13 * scale
    # the FP64 sums sit closer to the exact value than the compensated FP32 ones
    assert (np.abs(hyb - ic64)[wide].max() <= np.abs(plain - ic64)[wide].max())


def test_compensated_sum_is_exact_on_cancellation():
    # receiver 0 at 0 mV; neighbours: one at -30 mV (term +B), K at -1e-4 mV
    # (K identical small terms t), one at +30 mV (term -B, the exact negative
    # of +B since V -> -V negates the computed term exactly).  The true sum of
    # the computed terms is K*t; plain FP32 summation loses the small terms
    # against B, the compensated sum keeps them.
    K = 2000
    v = np.array([0.0, -30.0] + [-1e-4] * K + [30.0])
    n = len(v)
    col = np.arange(1, n, dtype=np.int32)
    w = np.full(n - 1, 0.06)
    rp = np.array([0, n - 1] + [n - 1] * (n - 1), np.int64)
    one = O.gap_currents(O.FP32, v[:3], np.array([0, 1, 1, 1], np.int64),
                         np.array([2], np.int32), np.array([0.06]))[0]
    exact = K * one            # exact in double
    plain = O.gap_currents(O.FP32, v, rp, col, w)[0]
    comp = O.gap_currents(O.FP32K, v, rp, col, w)[0]
    ulp = np.spacing(np.float32(exact))
    assert abs(comp - exact) <= 1 * ulp, (comp, exact)
    assert abs(plain - exact) > 8 * ulp, (plain, exact)


def test_synchrony_invariant():
    # identical cells with identical drive stay identical, bit for bit, on
    # any graph: every I_c is exactly 0
    rng = np.random.default_rng(2)
    n = 30
    conn = np.where(rng.random((n, n)) < 0.3, rng.uniform(0.01, 0.08, (n, n)), 0.0)
    np.fill_diagonal(conn, 0)
    rp, col, w = O.dense_to_csr(conn)
    for prec in (O.FP64, O.FP32, O.FP32K):
        st = O.initial_state(n)
        out, _ = O.run(prec, st, 600, rp, col, w, amp=6.0, t0=100, t1=300)
        one = O.initial_state(1)
        ref, _ = O.run(prec, one, 600, amp=6.0, t0=100, t1=300)
        assert (out == ref).all(), prec


@pytest.mark.parametrize("prec,delta,tol", [(O.FP64, 1e-4, 1e-8), (O.FP32, 1e-2, 2e-6), (O.FP32K, 1e-2, 2e-6)])
@pytest.mark.parametrize("var,v0,gate,gate_value", [(O.V_DEND, -8.5, O.CALCIUM_R, 1.0),
                                                    (O.V_SOMA, -25.0, O.POTASSIUM_X_S, 0.0),
                                                    (O.V_AXON, -25.0, O.POTASSIUM_X_A, 0.0)])
def test_removable_singularity_takes_its_limit(prec, delta, tol, var, v0, gate, gate_value):
    """beta_r = 0.02 x / (exp(x/5) - 1), x = V_dend + 8.5, and alpha_x =
    0.13 x / (1 - exp(-x/10)), x = V + 25, are 0/0 at x = 0 exactly (the
    reference computes NaN there and the network follows).  The oracle, like
    the GPU path, uses the limits 0.02 * 5 and 0.13 * 10: one step from the
    singular voltage gives the midpoint of the steps from V0 -/+ delta
    (the gate's update is smooth in V).  The gate value is chosen so the
    update is sensitive to the rate (dr/dbeta_r = -0.01 r, dx/dalpha_x =
    0.05 (1 - x)): a limit off by 10% moves the result by 1e-4 or more."""
    out = []
    for v in (v0 - delta, v0, v0 + delta):
        st = O.initial_state(1)
        st[var] = v
        st[gate] = gate_value
        O.run(prec, st, 1, trace=False)
        assert np.isfinite(st).all()
        out.append(st[gate, 0])
    assert abs(out[1] - 0.5 * (out[0] + out[2])) < tol, out


@pytest.mark.skipif(not os.path.exists(os.path.join(O.REF_DIR, "checker.x")),
                    reason="reference checker not built")
def test_reference_checker_accepts_oracle_trace(golden):
    # run the reference's own regression checker on the oracle's FP32 trace
    (rp, col, w), gcal = _regression_setup()
    st = O.initial_state(10)
    st[O.G_CAL] = gcal
    out, _ = O.run(O.FP32, st, 10000, rp, col, w, amp=6.0, t0=20000, t1=20499)
    z = golden("stable_reg10")
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "InferiorOlive_Output.txt"), "w") as fh:
            for row in out[z["step_index"]]:
                fh.write("".join("%.8f " % x for x in row) + "\n")
        with open(os.path.join(d, "InferiorOlive_Output_Baseline.txt"), "w") as fh:
            for row in z["trace"]:
                fh.write("".join("%.8f " % x for x in row) + "\n")
        shutil.copy(os.path.join(O.REF_DIR, "checker.x"), d)
        r = subprocess.run(["./checker.x", "10"], cwd=d, capture_output=True, text=True)
    assert "Check was passed successfully" in r.stdout, r.stdout
