"""Host side of the library (no GPU needed): exported symbols, configuration,
seeded generators, file formats and error paths."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

import infoli_amd as I
import oracle_lib as O

ROOT = O.ROOT
HEADER = os.path.join(ROOT, "include", "infoli_gpu.h")


def _has_gpu():
    try:
        I.Network(1).close()
        return True
    except I.InfoliError:
        return False


HAS_GPU = None


def has_gpu():
    global HAS_GPU
    if HAS_GPU is None:
        HAS_GPU = _has_gpu()
    return HAS_GPU


# ---------------------------------------------------------------- boundary

def test_library_exports_every_declared_function():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(infoli_[a-z0-9_]+)\s*\(", text))
    assert len(names) >= 30
    L = I.library()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", I.library_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\b(infoli_[a-z0-9_]+)\b", out))
    assert names <= exported, names - exported


def test_python_constants_match_header():
    """The binding's precision, launch-mode and gather codes and the FP64
    degree default are the values include/infoli_gpu.h defines."""
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "infoli_gpu.h")).read()
    defs = {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define INFOLI_([A-Z0-9_]+)\s+(-?\d+)", hdr)}
    names = ["FP64", "FP32", "FP32_KAHAN", "FP64_DEGREE_DEFAULT",
             "EXEC_AUTO", "EXEC_STEP", "EXEC_GRAPH", "EXEC_PERSISTENT",
             "GATHER_DIRECT", "GATHER_LDS", "GATHER_STAGED", "GATHER_TILED", "GATHER_SWEEP",
             "GATHER_NETS", "ERR_ARG", "ERR_STATE", "ERR_NONFINITE"]
    for name in names:
        assert getattr(I, name) == defs[name], name


def test_version_and_strerror():
    L = I.library()
    assert b"gfx950" in L.infoli_version()
    assert L.infoli_strerror(0) == b"ok"
    assert L.infoli_strerror(-5) == b"non-finite voltage"


# ---------------------------------------------------------------- config

def test_default_params_equal_reference_default_conf():
    d, f = O.read_conf()
    p = I.Params.default()
    assert (p.values == d).all()
    assert (p.values_f32 == f).all()


def test_read_config_reference_file():
    p = I.Params.from_file(os.path.join(O.GOLDEN, "default.conf"))
    d, f = O.read_conf()
    assert (p.values == d).all() and (p.values_f32 == f).all()


def test_read_config_errors(tmp_path):
    with pytest.raises(I.InfoliError) as e:
        I.Params.from_file(tmp_path / "missing.conf")
    assert e.value.code == -6
    lines = open(os.path.join(O.GOLDEN, "default.conf")).read().splitlines()
    bad = list(lines)
    bad[5] = "G_K_S = five"
    (tmp_path / "bad.conf").write_text("\n".join(bad) + "\n")
    with pytest.raises(I.InfoliError) as e:
        I.Params.from_file(tmp_path / "bad.conf")
    assert e.value.code == -7 and "line 6" in str(e.value) and "Improper Parameter" in str(e.value)
    (tmp_path / "short.conf").write_text("\n".join(lines[:30]) + "\n")
    with pytest.raises(I.InfoliError) as e:
        I.Params.from_file(tmp_path / "short.conf")
    assert e.value.code == -7 and "line 31" in str(e.value)
    # blank lines between entries are skipped, as fscanf does
    (tmp_path / "blank.conf").write_text("\n\n".join(lines) + "\n")
    assert (I.Params.from_file(tmp_path / "blank.conf").values == O.read_conf()[0]).all()


# ---------------------------------------------------------------- generators

M64 = (1 << 64) - 1


def _mix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def _bits(seed, stream, a, b):
    h = _mix64((seed + stream * 0xD1B54A32D192ED03) & M64)
    h = _mix64((h + a * 0x9E6C63D0676A9A99) & M64)
    return _mix64((h + b) & M64)


def _u01(x):
    return ((x >> 11) + 1) * 2.0 ** -53


def _skip_positions(seed, stream, row, length, p):
    if length <= 0 or not p > 0:
        return []
    if p >= 1:
        return list(range(length))
    lq = math.log1p(-p)
    out, x, k = [], -1, 0
    while True:
        u = _u01(_bits(seed, stream, row, k))
        k += 1
        skip = math.floor(math.log(u) / lq)
        if not skip < length - 1 - x:
            return out
        x += skip + 1
        out.append(x)


def _sym_reference(seed, n, p, lo, hi):
    """Dense symmetric matrix from the generator's definition, then CSR by
    scanning rows (an independent construction of the same graph)."""
    m = np.zeros((n, n))
    for i in range(n - 1):
        for x in _skip_positions(seed, 0, i, n - 1 - i, p):
            j = i + 1 + x
            wv = lo if hi == lo else lo + (hi - lo) * _u01(_bits(seed, 1, i, j))
            m[i, j] = m[j, i] = wv
    return O.dense_to_csr(m)


@pytest.mark.parametrize("n,p", [(1, 0.5), (2, 1.0), (50, 1.0), (300, 0.05), (257, 0.3), (400, 0.0)])
def test_gen_symmetric_equals_definition(n, p):
    seed = 12345
    rp, col, w = I.gen_symmetric(seed, n, p, 0.02, 0.06)
    rp2, col2, w2 = _sym_reference(seed, n, p, 0.02, 0.06)
    assert (rp == rp2).all() and (col == col2).all() and (w == w2).all()


def test_gen_symmetric_properties():
    n, p = 3000, 0.02
    rp, col, w = I.gen_symmetric(99, n, p, 0.02, 0.06)
    deg = np.diff(rp)
    rows = np.repeat(np.arange(n), deg)
    assert (col != rows).all()                       # no self junctions
    for i in range(0, n, 97):                        # rows strictly ascending
        assert (np.diff(col[rp[i]:rp[i + 1]]) > 0).all()
    a = {(int(r), int(c)): float(x) for r, c, x in zip(rows, col, w)}
    assert all(a[(c, r)] == x for (r, c), x in a.items())   # symmetric, w_ij = w_ji
    mu, var = p * (n - 1), p * (1 - p) * (n - 1)
    assert abs(deg.mean() - mu) < 5 * math.sqrt(var / n)
    assert 0.8 * var < deg.var() < 1.25 * var
    assert w.min() >= 0.02 and w.max() <= 0.06
    # determinism and seed dependence
    rp2, col2, w2 = I.gen_symmetric(99, n, p, 0.02, 0.06)
    assert (col == col2).all() and (w == w2).all()
    rp3, col3, _ = I.gen_symmetric(100, n, p, 0.02, 0.06)
    assert len(col3) != len(col) or (col3 != col).any()


def test_gen_directed_reference_semantics():
    # p = 1: every other cell, ascending, conductance w0 (the reference's
    # generator always connects at p = 1 since rand()/RAND_MAX <= 1)
    rp, col, w = I.gen_directed(1, 10, 1.0, 0.04)
    assert (np.diff(rp) == 9).all()
    for i in range(10):
        assert list(col[rp[i]:rp[i + 1]]) == [j for j in range(10) if j != i]
    assert (w == 0.04).all()
    rp, col, w = I.gen_directed(1, 10, 0.0, 0.04)
    assert rp[-1] == 0
    # statistics and the per-receiver definition
    n, p = 500, 0.1
    rp, col, w = I.gen_directed(7, n, p, 0.04)
    for i in (0, 17, 499):
        xs = _skip_positions(7, 6, i, n - 1, p)
        assert list(col[rp[i]:rp[i + 1]]) == [x if x < i else x + 1 for x in xs]
    assert abs(np.diff(rp).mean() - p * (n - 1)) < 5 * math.sqrt(p * (1 - p) * (n - 1) / n)


def test_gen_values():
    u = I.gen_uniform(3, 2, 20000, 0.5, 1.1)
    assert u.min() >= 0.5 and u.max() <= 1.1 and abs(u.mean() - 0.8) < 0.01
    assert u[5] == 0.5 + 0.6 * _u01(_bits(3, 2, 5, 0))
    z = I.gen_normal(3, 3, 20000, 1.0, 0.5)
    assert abs(z.mean() - 1.0) < 0.02 and abs(z.std() - 0.5) < 0.02
    m = I.gen_mask(3, 4, 20000, 0.2)
    assert abs(m.mean() - 0.2) < 0.015 and set(np.unique(m)) <= {0, 1}
    assert (I.gen_uniform(3, 2, 100, 0.5, 1.1) == u[:100]).all()


def test_gen_rand_init_equals_definition():
    """RAND_INIT's g_CaL 0.6 + 0.2*(k % 100)/100 and odd warm-up step counts
    below ceil(100/DELTA) (src/single_threaded/infoli.c:212-222), k and u the
    cell's draws of stream 7."""
    n, seed = 500, 99
    g, st = I.gen_rand_init(seed, n, 0.05)
    for i in range(n):
        k, u = _bits(seed, 7, i, 0), _bits(seed, 7, i, 1)
        assert g[i] == 0.6 + (0.2 * (k % 100)) / 100
        assert st[i] == (u % 2000) | 1
    assert (st % 2 == 1).all() and st.min() >= 1 and st.max() < 2000
    assert g.min() >= 0.6 and g.max() <= 0.6 + 0.2 * 99 / 100 + 1e-15
    g2, st2 = I.gen_rand_init(seed + 1, n, 0.05)
    assert not np.array_equal(st, st2)


def test_gen_argument_errors():
    with pytest.raises(I.InfoliError):
        I.gen_symmetric(1, 0, 0.1, 0.04)
    with pytest.raises(I.InfoliError):
        I.gen_symmetric(1, 10, 1.5, 0.04)
    with pytest.raises(I.InfoliError):
        I.gen_directed(1, 10, float("nan"), 0.04)


# ---------------------------------------------------------------- files

def test_read_conn_dense_reference_fixture():
    path = os.path.join(O.GOLDEN, "cellConnections.txt")
    rp, col, w = I.read_conn_dense(path, 10)
    m = np.loadtxt(path)
    rp2, col2, w2 = O.dense_to_csr(m)
    assert (rp == rp2).all() and (col == col2).all() and (w == w2).all()
    assert rp[-1] == 90


def test_read_conn_dense_threshold_and_short(tmp_path):
    m = np.array([[0, 5e-8, 0.04], [0.02, 0, 0], [0, -0.01, 0]])
    f = tmp_path / "c.txt"
    f.write_text("\n".join(" ".join(repr(float(x)) for x in r) for r in m) + "\n")
    rp, col, w = I.read_conn_dense(f, 3)                 # ST rule: != 0
    assert list(np.diff(rp)) == [1, 2, 1]
    rp, col, w = I.read_conn_dense(f, 3, threshold=1e-7)  # stable_version rule
    assert list(np.diff(rp)) == [1, 0, 1]
    f.write_text("0 0.04\n0.04\n")
    with pytest.raises(I.InfoliError) as e:
        I.read_conn_dense(f, 2 + 1)
    assert e.value.code == -6


def test_read_conn_triplets(tmp_path):
    f = tmp_path / "t.txt"
    f.write_text("1 0 0.04\n2 0 0.03\n0 1 0.05\n2 1 0.0\n0 2 0.02\n")
    rp, col, w = I.read_conn_triplets(f, 3)
    assert list(rp) == [0, 2, 3, 4]
    assert list(col) == [1, 2, 0, 0] and list(w) == [0.04, 0.03, 0.05, 0.02]
    f.write_text("1 0 0.04\n5 0 0.03\n")
    with pytest.raises(I.InfoliError):
        I.read_conn_triplets(f, 3)


def test_read_values_and_iapp_file(tmp_path):
    g = I.read_values(os.path.join(O.GOLDEN, "gcal_file.txt"), 10)
    assert list(g) == [0.68, 0.76, 0.689, 0.67, 0.7, 0.775, 0.685, 0.702, 0.696, 0.67]
    with pytest.raises(I.InfoliError):
        I.read_values(os.path.join(O.GOLDEN, "gcal_file.txt"), 11)
    f = tmp_path / "iapp.txt"
    f.write_text("6.00 0.00 1.5\n0 0 0 7\n1 2 3\n")
    tab = I.read_iapp_file(f, 3)
    assert tab.shape == (3, 3) and list(tab[0]) == [6.0, 0.0, 1.5] and list(tab[1]) == [0, 0, 0]
    f.write_text("6.00 0.00 1.5\n0 0\n")
    with pytest.raises(I.InfoliError) as e:
        I.read_iapp_file(f, 3)
    assert "enough elements" in str(e.value)


# ---------------------------------------------------------------- no-GPU behaviour

def test_network_without_gpu_fails_loudly():
    if has_gpu():
        pytest.skip("a GPU is present")
    with pytest.raises(I.InfoliError) as e:
        I.Network(10)
    assert e.value.code == -9


def test_driver_usage_and_failure():
    exe = os.path.join(ROOT, "infoli_amd", "bin", "infoli.x")
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 0 and "-n denotes number of neurons" in r.stdout
    if not has_gpu():
        r = subprocess.run([exe, "-n", "4", "-t", "1"], capture_output=True, text=True,
                           cwd=os.path.dirname(exe))
        assert r.returncode != 0 and "no HIP device" in r.stdout
