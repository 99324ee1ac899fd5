"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module; the product (infoli_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

FP64, FP32, FP32K = 0, 1, 2
NVARS = 19
(V_DEND, CALCIUM_R, POTASSIUM_S, HCURRENT_Q, CA2PLUS, I_CAH, G_CAL, V_SOMA,
 SODIUM_M, SODIUM_H, POTASSIUM_N, POTASSIUM_P, POTASSIUM_X_S, CALCIUM_K,
 CALCIUM_L, V_AXON, SODIUM_M_A, SODIUM_H_A, POTASSIUM_X_A) = range(NVARS)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def read_conf(path=None):
    """(double values, float32 values) of a 42-line "NAME = value" file
    (default: the reference's bin/default.conf kept as a fixture)."""
    path = path or os.path.join(GOLDEN, "default.conf")
    vals = []
    with open(path) as fh:
        for line in fh:
            if line.strip():
                vals.append(line.split("=")[1].strip())
    assert len(vals) == 42, path
    d = np.array([float(v) for v in vals], dtype=np.float64)
    f = np.array([np.float32(v) for v in vals], dtype=np.float32)
    return d, f


def default_conf():
    return read_conf()


def initial_state(n, conf=None):
    """[19, n] float64 state at the configured initial values."""
    d, _ = default_conf() if conf is None else conf
    return np.repeat(d[23:42, None], n, axis=1).astype(np.float64).copy()


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.oracle_run.argtypes = [ctypes.c_int, P, P, i64, P, P, P, P, P,
                                 ctypes.c_double, i64, i64, i64, i64, P, P, i64,
                                 i64, i64, P, ctypes.c_int, P]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_gap_currents.argtypes = [ctypes.c_int, i64, P, P, P, P, P]
        L.oracle_gap_currents.restype = ctypes.c_int
        L.oracle_max_threads.restype = ctypes.c_int
        L.oracle_set_fp64_degree.argtypes = [i64]
        L.oracle_set_fp64_degree.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def csr_empty(n):
    return (np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float64))


def run(prec, state, nsteps, row_ptr=None, col=None, w=None, start_step=0,
        stride=1, iapp_base=None, mask=None, amp=0.0, t0=0, t1=0,
        table=None, table_t0=0, conf=None, nthreads=0, trace=True):
    """Advance `state` ([19, n] float64, modified in place) by nsteps.

    Returns (trace [records, n] or None, elapsed seconds of the step loop).
    """
    n = state.shape[1]
    assert state.shape == (NVARS, n) and state.dtype == np.float64
    assert state.flags.c_contiguous
    if row_ptr is None:
        row_ptr, col, w = csr_empty(n)
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    w = np.ascontiguousarray(w, np.float64)
    assert row_ptr.shape == (n + 1,) and row_ptr[-1] == len(col) == len(w)
    cd, cf = default_conf() if conf is None else conf
    cd = np.ascontiguousarray(cd, np.float64)
    cf = np.ascontiguousarray(cf, np.float32)
    if iapp_base is not None:
        iapp_base = np.ascontiguousarray(iapp_base, np.float64)
        assert iapp_base.shape == (n,)
    if mask is not None:
        mask = np.ascontiguousarray(mask, np.uint8)
        assert mask.shape == (n,)
    tab_n = 0
    if table is not None:
        table = np.ascontiguousarray(table, np.float64)
        assert table.ndim == 2 and table.shape[1] == n
        tab_n = table.shape[0]
    tr = None
    if trace:
        nrec = sum(1 for t in range(start_step, start_step + nsteps) if t % stride == 0)
        tr = np.zeros((nrec, n), np.float64)
    el = ctypes.c_double(0)
    rc = lib().oracle_run(prec, _ptr(cd), _ptr(cf), n, _ptr(row_ptr), _ptr(col),
                          _ptr(w), _ptr(iapp_base), _ptr(mask), float(amp),
                          int(t0), int(t1), int(table_t0), tab_n, _ptr(table),
                          _ptr(state), int(start_step), int(nsteps), int(stride),
                          _ptr(tr), int(nthreads), ctypes.byref(el))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed: {rc}")
    return tr, el.value


def gap_currents(prec, vdend, row_ptr, col, w):
    n = len(vdend)
    vdend = np.ascontiguousarray(vdend, np.float64)
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    w = np.ascontiguousarray(w, np.float64)
    ic = np.zeros(n, np.float64)
    rc = lib().oracle_gap_currents(prec, n, _ptr(row_ptr), _ptr(col), _ptr(w),
                                   _ptr(vdend), _ptr(ic))
    if rc != 0:
        raise RuntimeError(f"oracle_gap_currents failed: {rc}")
    return ic


def dense_to_csr(mat, threshold=0.0):
    """Receiver-major CSR from a dense [sender, receiver] conductance matrix,
    senders ascending (the reading order of the reference's dense file)."""
    m = np.asarray(mat, np.float64)
    n = m.shape[0]
    rows, cols, ws = [0], [], []
    for rcv in range(n):
        for snd in range(n):
            v = m[snd, rcv]
            if (v != 0.0) if threshold == 0.0 else (v >= threshold):
                cols.append(snd)
                ws.append(v)
        rows.append(len(cols))
    return (np.array(rows, np.int64), np.array(cols, np.int32), np.array(ws, np.float64))


def set_fp64_degree(d):
    """FP32K cells with at least d neighbours sum FP64 terms (0 = none)."""
    lib().oracle_set_fp64_degree(int(d))
