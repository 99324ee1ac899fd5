"""ctypes binding of libinfoli_gpu.so and the Network class.

Every array crossing the boundary is a contiguous numpy array; the library
copies inputs, so callers keep ownership.  Errors come back as InfoliError
carrying the library's status code and message (the reference printed and
exit()ed instead, e.g. src/utils/utils.c:143-153).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libinfoli_gpu.so")

FP64, FP32, FP32_KAHAN = 0, 1, 2
FP64_DEGREE_DEFAULT = 1280   # FP32_KAHAN: cells of at least this degree sum FP64 terms
_PREC_NAMES = {"fp64": FP64, "fp32": FP32, "fp32k": FP32_KAHAN, "fp32_kahan": FP32_KAHAN}

EXEC_AUTO, EXEC_STEP, EXEC_GRAPH, EXEC_PERSISTENT = 0, 1, 2, 3
GATHER_DIRECT, GATHER_LDS, GATHER_STAGED, GATHER_TILED, GATHER_SWEEP, GATHER_NETS = 0, 1, 2, 3, 4, 5

# per-cell variables, bin/default.conf lines 24-42 order
VARS = ("V_dend", "Calcium_r", "Potassium_s", "Hcurrent_q", "Ca2Plus", "I_CaH",
        "g_CaL", "V_soma", "Sodium_m", "Sodium_h", "Potassium_n", "Potassium_p",
        "Potassium_x_s", "Calcium_k", "Calcium_l", "V_axon", "Sodium_m_a",
        "Sodium_h_a", "Potassium_x_a")
_VAR_INDEX = {v: i for i, v in enumerate(VARS)}

ERR_ARG, ERR_STATE, ERR_NONFINITE = -1, -4, -5


class InfoliError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"infoli error {code}: {msg}")
        self.code = code


class _Params(ctypes.Structure):
    _fields_ = [("v", ctypes.c_double * 42), ("f", ctypes.c_float * 42)]


class _Exec(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("lanes_per_cell", ctypes.c_int),
                ("block", ctypes.c_int), ("steps_per_launch", ctypes.c_int),
                ("gather", ctypes.c_int), ("lds_kb", ctypes.c_int),
                ("tile_cells", ctypes.c_int), ("ode_split", ctypes.c_int),
                ("fast_math", ctypes.c_int)]


_lib = None


def library_path():
    return _LIB_PATH


def library():
    """Load libinfoli_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"{_LIB_PATH} not built: run `make -C infoli_amd/csrc` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(_LIB_PATH)
    P, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    u64 = ctypes.c_uint64
    sig = {
        "infoli_params_default": ([P], i32),
        "infoli_read_config": ([ctypes.c_char_p, P], i32),
        "infoli_create": ([P, P, i64, i32, i32], i32),
        "infoli_destroy": ([P], None),
        "infoli_set_connectivity_csr": ([P, P, P, P], i32),
        "infoli_set_partition": ([P, i32, P], i32),
        "infoli_set_fp64_degree": ([P, i64], i32),
        "infoli_write_state": ([P, ctypes.c_char_p], i32),
        "infoli_save_checkpoint": ([P, ctypes.c_char_p], i32),
        "infoli_load_checkpoint": ([P, ctypes.c_char_p], i32),
        "infoli_isolated_steps": ([P, P], i32),
        "infoli_gen_rand_init": ([u64, i64, d, P, P], i32),
        "infoli_get_fp64_degree": ([P], i64),
        "infoli_set_gcal": ([P, P], i32),
        "infoli_set_iapp": ([P, P], i32),
        "infoli_set_iapp_pulse": ([P, P, d, i64, i64], i32),
        "infoli_set_iapp_table": ([P, i64, i64, P], i32),
        "infoli_set_state": ([P, i32, P], i32),
        "infoli_get_state": ([P, i32, P], i32),
        "infoli_advance": ([P, i64, i64, P], i32),
        "infoli_trace_records": ([P, i64, i64], i64),
        "infoli_prepare": ([P, i64, i64], i32),
        "infoli_last_launches": ([P, P, P, P], i32),
        "infoli_step_count": ([P], i64),
        "infoli_gap_currents": ([P, P], i32),
        "infoli_gap_currents_noalloc": ([P, P], i32),
        "infoli_get_exec": ([P, P], i32),
        "infoli_set_exec": ([P, P], i32),
        "infoli_autotune": ([P, i64, P], i32),
        "infoli_last_timing": ([P, P, P, P], i32),
        "infoli_set_profiling": ([P, i32], i32),
        "infoli_synchronize": ([P], i32),
        "infoli_gen_symmetric": ([u64, i64, d, d, d, P, P, P, P], i32),
        "infoli_gen_directed": ([u64, i64, d, d, P, P, P, P], i32),
        "infoli_gen_uniform": ([u64, i32, i64, d, d, P], i32),
        "infoli_gen_normal": ([u64, i32, i64, d, d, P], i32),
        "infoli_gen_mask": ([u64, i32, i64, d, P], i32),
        "infoli_read_conn_dense": ([ctypes.c_char_p, i64, d, P, P, P, P], i32),
        "infoli_read_conn_triplets": ([ctypes.c_char_p, i64, d, P, P, P, P], i32),
        "infoli_read_values": ([ctypes.c_char_p, i64, P], i32),
        "infoli_read_iapp_file": ([ctypes.c_char_p, i64, i64, P, P], i32),
        "infoli_tiled_layout_check": ([i64, P, P, P, i32, i32, i32, i32, i32, i32, i32, i64], i32),
        "infoli_last_error": ([], ctypes.c_char_p),
        "infoli_strerror": ([i32], ctypes.c_char_p),
        "infoli_version": ([], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _check(rc, what=""):
    if rc != 0:
        L = library()
        msg = L.infoli_last_error().decode() or L.infoli_strerror(rc).decode()
        raise InfoliError(rc, f"{what}: {msg}" if what else msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _prec(p):
    if isinstance(p, str):
        return _PREC_NAMES[p.lower()]
    return int(p)


def _var(v):
    return _VAR_INDEX[v] if isinstance(v, str) else int(v)


class Params:
    """The 42 configuration values (bin/default.conf order), each as the
    file's double and as the reference's float reading."""

    def __init__(self, raw=None):
        self._raw = raw if raw is not None else _Params()
        if raw is None:
            _check(library().infoli_params_default(ctypes.byref(self._raw)), "params_default")

    @classmethod
    def default(cls):
        return cls()

    @classmethod
    def from_file(cls, path):
        raw = _Params()
        _check(library().infoli_read_config(os.fsencode(path), ctypes.byref(raw)), "read_config")
        return cls(raw)

    @property
    def values(self):
        return np.array(self._raw.v[:], np.float64)

    @property
    def values_f32(self):
        return np.array(self._raw.f[:], np.float32)

    def set(self, index, value):
        self._raw.v[index] = float(value)
        self._raw.f[index] = float(np.float32(value))


class Network:
    """One network (or a block-diagonal batch of networks) resident on a GPU.

    The per-timestep update of the reference -- gap-junction loop
    src/openmp/infoli.c:496-520 then compartment loop :548-679 -- runs in the
    HIP library on `advance`.
    """

    def __init__(self, n_cells, precision="fp32k", params=None, device=0):
        L = library()
        self.n = int(n_cells)
        self.precision = _prec(precision)
        self.params = params if params is not None else Params()
        h = ctypes.c_void_p()
        _check(L.infoli_create(ctypes.byref(h), ctypes.byref(self.params._raw), self.n,
                               self.precision, int(device)), "create")
        self._h = h

    # lifecycle
    def close(self):
        if getattr(self, "_h", None):
            library().infoli_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # inputs
    def set_connectivity(self, row_ptr, col, w):
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        w = np.ascontiguousarray(w, np.float64)
        if row_ptr.shape != (self.n + 1,):
            raise ValueError(f"row_ptr must have {self.n + 1} entries")
        if len(col) != row_ptr[-1] or len(w) != row_ptr[-1]:
            raise ValueError("col and w must have row_ptr[-1] entries")
        _check(library().infoli_set_connectivity_csr(self._h, _p(row_ptr), _p(col), _p(w)),
               "set_connectivity")

    def set_partition(self, offsets):
        offsets = np.ascontiguousarray(offsets, np.int64)
        _check(library().infoli_set_partition(self._h, len(offsets) - 1, _p(offsets)),
               "set_partition")

    def write_state(self, path):
        """printState's 16-column dump of every cell (%.8lf)."""
        _check(library().infoli_write_state(self._h, os.fsencode(path)), "write_state")

    def save_checkpoint(self, path):
        """All 19 variables (FP64) and the step count, binary."""
        _check(library().infoli_save_checkpoint(self._h, os.fsencode(path)), "save_checkpoint")

    def load_checkpoint(self, path):
        _check(library().infoli_load_checkpoint(self._h, os.fsencode(path)), "load_checkpoint")

    def isolated_steps(self, steps):
        """Every cell alone (no gap-junction or applied current) for steps[i]
        steps: RAND_INIT's warm-up.  The step count does not change."""
        steps = np.ascontiguousarray(steps, np.int64)
        if steps.shape != (self.n,):
            raise ValueError(f"steps must have {self.n} entries")
        _check(library().infoli_isolated_steps(self._h, _p(steps)), "isolated_steps")

    def set_fp64_degree(self, min_degree):
        """FP32 compensated only: cells with at least `min_degree` neighbours
        compute and sum their gap-junction terms in FP64 (0 = none)."""
        _check(library().infoli_set_fp64_degree(self._h, int(min_degree)), "set_fp64_degree")

    @property
    def fp64_degree(self):
        return int(library().infoli_get_fp64_degree(self._h))

    def _cells(self, values, dtype=np.float64):
        a = np.ascontiguousarray(values, dtype)
        if a.shape != (self.n,):
            raise ValueError(f"expected {self.n} values, got shape {a.shape}")
        return a

    def set_gcal(self, values):
        _check(library().infoli_set_gcal(self._h, _p(self._cells(values))), "set_gcal")

    def set_iapp(self, base=None):
        b = None if base is None else self._cells(base)
        _check(library().infoli_set_iapp(self._h, _p(b)), "set_iapp")

    def set_iapp_pulse(self, amp, t0, t1, mask=None):
        m = None if mask is None else self._cells(mask, np.uint8)
        _check(library().infoli_set_iapp_pulse(self._h, _p(m), float(amp), int(t0), int(t1)),
               "set_iapp_pulse")

    def set_iapp_table(self, t_first, table):
        if table is None:
            _check(library().infoli_set_iapp_table(self._h, 0, 0, None), "set_iapp_table")
            return
        tab = np.ascontiguousarray(table, np.float64)
        if tab.ndim != 2 or tab.shape[1] != self.n:
            raise ValueError("table must be [steps, n_cells]")
        _check(library().infoli_set_iapp_table(self._h, int(t_first), tab.shape[0], _p(tab)),
               "set_iapp_table")

    def set_state(self, var, values):
        _check(library().infoli_set_state(self._h, _var(var), _p(self._cells(values))), "set_state")

    def get_state(self, var):
        out = np.zeros(self.n, np.float64)
        _check(library().infoli_get_state(self._h, _var(var), _p(out)), "get_state")
        return out

    def state(self):
        """[19, n] array of every per-cell variable."""
        return np.stack([self.get_state(i) for i in range(len(VARS))])

    # stepping
    @property
    def step_count(self):
        return library().infoli_step_count(self._h)

    def trace_records(self, n_steps, stride):
        return library().infoli_trace_records(self._h, int(n_steps), int(stride))

    def advance(self, n_steps, trace_stride=None, allow_nonfinite=False):
        """Advance n_steps; returns the V_axon trace [records, n] (records
        after steps t with t % trace_stride == 0) or None."""
        L = library()
        tr = None
        if trace_stride:
            nrec = self.trace_records(n_steps, trace_stride)
            tr = np.zeros((nrec, self.n), np.float64)
        rc = L.infoli_advance(self._h, int(n_steps), int(trace_stride or 0), _p(tr))
        if rc != 0 and not (allow_nonfinite and rc == ERR_NONFINITE):
            _check(rc, "advance")
        return tr

    def prepare(self, n_steps, trace_stride=None):
        """Build the layout, trace ring and graph a following
        advance(n_steps, trace_stride) uses, without running a step."""
        _check(library().infoli_prepare(self._h, int(n_steps), int(trace_stride or 0)), "prepare")

    def last_launches(self):
        """How the last advance launched: graph replays, steps launched one
        by one, graphs captured during the call."""
        r, d, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(library().infoli_last_launches(self._h, ctypes.byref(r), ctypes.byref(d), ctypes.byref(c)),
               "last_launches")
        return dict(graph_replays=r.value, direct_steps=d.value, captures=c.value)

    def gap_currents(self, noalloc=False):
        """I_c of every cell from the current V_dend.  noalloc: through the
        entry point that allocates nothing (static device scratch, at most
        16 MiB of values)."""
        out = np.zeros(self.n, np.float64)
        if noalloc:
            _check(library().infoli_gap_currents_noalloc(self._h, _p(out)), "gap_currents_noalloc")
        else:
            _check(library().infoli_gap_currents(self._h, _p(out)), "gap_currents")
        return out

    def synchronize(self):
        _check(library().infoli_synchronize(self._h), "synchronize")

    # execution strategy
    @property
    def exec(self):
        e = _Exec()
        _check(library().infoli_get_exec(self._h, ctypes.byref(e)), "get_exec")
        return dict(mode=e.mode, lanes_per_cell=e.lanes_per_cell, block=e.block,
                    steps_per_launch=e.steps_per_launch, gather=e.gather, lds_kb=e.lds_kb,
                    tile_cells=e.tile_cells, ode_split=e.ode_split, fast_math=e.fast_math)

    def set_exec(self, mode=None, lanes_per_cell=None, block=None, steps_per_launch=None,
                 gather=None, lds_kb=None, tile_cells=None, ode_split=None, fast_math=None):
        cur = self.exec
        if gather is not None and gather != GATHER_NETS and mode is None and cur["gather"] == GATHER_NETS:
            # leaving whole networks per workgroup (persistent by construction):
            # the other gathers start from graph launches
            cur["mode"] = EXEC_GRAPH
        for k, v in (("mode", mode), ("lanes_per_cell", lanes_per_cell), ("block", block),
                     ("steps_per_launch", steps_per_launch), ("gather", gather),
                     ("lds_kb", lds_kb), ("tile_cells", tile_cells), ("ode_split", ode_split),
                     ("fast_math", fast_math)):
            if v is not None:
                cur[k] = int(v)
        e = _Exec(**cur)
        _check(library().infoli_set_exec(self._h, ctypes.byref(e)), "set_exec")

    def autotune(self, trial_steps=200):
        e = _Exec()
        _check(library().infoli_autotune(self._h, int(trial_steps), ctypes.byref(e)), "autotune")
        return self.exec

    def set_profiling(self, on=True):
        _check(library().infoli_set_profiling(self._h, int(bool(on))), "set_profiling")

    def last_timing(self):
        a, k = ctypes.c_double(), ctypes.c_double()
        nl = ctypes.c_int64()
        _check(library().infoli_last_timing(self._h, ctypes.byref(a), ctypes.byref(k),
                                            ctypes.byref(nl)), "last_timing")
        return dict(advance_ms=a.value, kernel_ms=k.value, kernel_launches=nl.value)


# ---- generators and file readers (host code of the library; no GPU needed) ----

def _two_call(fn, n, *args):
    row_ptr = np.zeros(n + 1, np.int64)
    nnz = ctypes.c_int64()
    _check(fn(*args, _p(row_ptr), None, None, ctypes.byref(nnz)))
    col = np.zeros(max(nnz.value, 1), np.int32)
    w = np.zeros(max(nnz.value, 1), np.float64)
    _check(fn(*args, _p(row_ptr), _p(col), _p(w), ctypes.byref(nnz)))
    return row_ptr, col[:nnz.value].copy(), w[:nnz.value].copy()


def gen_symmetric(seed, n, p, w_lo, w_hi=None):
    """Symmetric G(n, p) with per-junction w ~ U[w_lo, w_hi] (w_hi None: constant)."""
    w_hi = w_lo if w_hi is None else w_hi
    L = library()
    return _two_call(L.infoli_gen_symmetric, n, int(seed), int(n), float(p), float(w_lo),
                     float(w_hi))


def tiled_layout_check(row_ptr, col, w, precision="fp32k", gather=GATHER_TILED, lanes_per_cell=1,
                       tile_cells=512, lds_kb=2, ode_split=0, block=1024, nsa=1):
    """Host-only check of the tiled layout the library builds for this
    network and shape, and of one step's scratch reads and writes; raises
    InfoliError naming the first violation."""
    row_ptr = np.ascontiguousarray(row_ptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    w = np.ascontiguousarray(w, np.float64)
    n = len(row_ptr) - 1
    _check(library().infoli_tiled_layout_check(n, _p(row_ptr), _p(col), _p(w), _prec(precision), int(gather),
                                               int(lanes_per_cell), int(tile_cells), int(lds_kb),
                                               int(ode_split), int(block), int(nsa)), "tiled_layout_check")


def gen_directed(seed, n, p, w0):
    """The reference generator's semantics (src/openmp/infoli.c:337-375), seeded."""
    L = library()
    return _two_call(L.infoli_gen_directed, n, int(seed), int(n), float(p), float(w0))


def gen_uniform(seed, stream, n, lo, hi):
    out = np.zeros(n, np.float64)
    _check(library().infoli_gen_uniform(int(seed), int(stream), int(n), float(lo), float(hi),
                                        _p(out)))
    return out


def gen_rand_init(seed, n, delta=0.05):
    """RAND_INIT's per-cell g_CaL and warm-up step counts, seeded."""
    g = np.zeros(n, np.float64)
    st = np.zeros(n, np.int64)
    _check(library().infoli_gen_rand_init(int(seed), int(n), float(delta), _p(g), _p(st)), "gen_rand_init")
    return g, st


def gen_normal(seed, stream, n, mean, sd):
    out = np.zeros(n, np.float64)
    _check(library().infoli_gen_normal(int(seed), int(stream), int(n), float(mean), float(sd),
                                       _p(out)))
    return out


def gen_mask(seed, stream, n, frac):
    out = np.zeros(n, np.uint8)
    _check(library().infoli_gen_mask(int(seed), int(stream), int(n), float(frac), _p(out)))
    return out


def read_conn_dense(path, n, threshold=0.0):
    L = library()
    return _two_call(L.infoli_read_conn_dense, n, os.fsencode(path), int(n), float(threshold))


def read_conn_triplets(path, n, threshold=0.0):
    L = library()
    return _two_call(L.infoli_read_conn_triplets, n, os.fsencode(path), int(n), float(threshold))


def read_values(path, n):
    out = np.zeros(n, np.float64)
    _check(library().infoli_read_values(os.fsencode(path), int(n), _p(out)), "read_values")
    return out


def read_iapp_file(path, n, max_steps=1 << 40):
    L = library()
    cnt = ctypes.c_int64()
    _check(L.infoli_read_iapp_file(os.fsencode(path), int(n), int(max_steps), None,
                                   ctypes.byref(cnt)), "read_iapp_file")
    tab = np.zeros((cnt.value, n), np.float64)
    _check(L.infoli_read_iapp_file(os.fsencode(path), int(n), cnt.value, _p(tab),
                                   ctypes.byref(cnt)), "read_iapp_file")
    return tab
