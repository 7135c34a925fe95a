"""ctypes binding of the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcData(C.Structure):
    _fields_ = [("cols", C.c_void_p), ("stride", C.c_int64), ("n_vars", C.c_int),
                ("n_rows", C.c_int64), ("card", C.c_void_p)]


class OrcResult(C.Structure):
    _fields_ = [("g2", C.c_double), ("p_value", C.c_double), ("df", C.c_int32), ("indep", C.c_int32)]


class OrcStats(C.Structure):
    _fields_ = [("tests_per_level", C.c_int64 * 64), ("edges_after_level", C.c_int64 * 64),
                ("n_levels", C.c_int32), ("num_ci_test", C.c_int64), ("n_sepsets", C.c_int64),
                ("sep_z_total", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_table_cells.restype = C.c_int64
        L.orc_table_cells.argtypes = [C.POINTER(OrcData), C.c_int, C.c_int, C.c_void_p, C.c_int]
        L.orc_count.restype = None
        L.orc_count.argtypes = [C.POINTER(OrcData), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_g2_table.restype = None
        L.orc_g2_table.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int64,
                                   C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.orc_pchisq.restype = C.c_double
        L.orc_pchisq.argtypes = [C.c_double, C.c_int32]
        L.orc_chisq_crit.restype = C.c_double
        L.orc_chisq_crit.argtypes = [C.c_int32, C.c_double]
        L.orc_ci_test.restype = OrcResult
        L.orc_ci_test.argtypes = [C.POINTER(OrcData), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_double]
        L.orc_ci_batch.restype = C.c_int
        L.orc_ci_batch.argtypes = [C.POINTER(OrcData), C.c_void_p, C.c_int64, C.c_int, C.c_double,
                                   C.c_int, C.c_void_p]
        L.orc_choose.restype = C.c_int64
        L.orc_choose.argtypes = [C.c_int, C.c_int]
        L.orc_enumerate.restype = C.c_int64
        L.orc_enumerate.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int64]
        L.orc_unrank.restype = None
        L.orc_unrank.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_void_p]
        L.orc_pc_skeleton.restype = C.c_int
        L.orc_pc_skeleton.argtypes = [C.POINTER(OrcData), C.c_double, C.c_int, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                      C.POINTER(OrcStats)]
        L.orc_level0_sample.restype = C.c_int64
        L.orc_level0_sample.argtypes = [C.POINTER(OrcData), C.c_double, C.c_int64, C.c_int, C.c_void_p]
        L.orc_level_sample_ref128.restype = C.c_int64
        L.orc_level_sample_ref128.argtypes = [C.POINTER(OrcData), C.c_double, C.c_int, C.c_void_p, C.c_int64, C.c_int,
                                              C.POINTER(C.c_int64)]
        L.orc_level_sample.restype = C.c_int64
        L.orc_level_sample.argtypes = [C.POINTER(OrcData), C.c_double, C.c_int, C.c_void_p, C.c_int64, C.c_int,
                                       C.POINTER(C.c_int64)]
        L.orc_mle.restype = C.c_int64
        L.orc_mle.argtypes = [C.POINTER(OrcData), C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_void_p]
        _LIB = L
    return _LIB


class Data:
    """Holds the numpy buffers alive for the C struct."""

    def __init__(self, cols: np.ndarray, n_rows: int, card: np.ndarray):
        self.cols = np.ascontiguousarray(cols, dtype=np.uint8)
        self.card = np.ascontiguousarray(card, dtype=np.int32)
        self.n_rows = int(n_rows)
        self.s = OrcData(self.cols.ctypes.data, self.cols.shape[1], self.cols.shape[0],
                         self.n_rows, self.card.ctypes.data)

    @classmethod
    def of(cls, sample):
        return cls(sample.cols, sample.n_rows, sample.card)

    @property
    def ref(self):
        return C.byref(self.s)


def _z(z):
    return np.ascontiguousarray(np.asarray(z if z is not None else [], dtype=np.int32))


def counts(data: Data, x, y, z=()):
    z = _z(z)
    cells = lib().orc_table_cells(data.ref, x, y, z.ctypes.data, len(z))
    out = np.zeros(cells, dtype=np.int32)
    lib().orc_count(data.ref, x, y, z.ctypes.data, len(z), out.ctypes.data)
    dimz = int(np.prod(data.card[z])) if len(z) else 1
    return out.reshape(dimz, int(data.card[x]), int(data.card[y]))


def g2_table(table: np.ndarray):
    t = np.ascontiguousarray(table, dtype=np.int32)
    g2, df = C.c_double(), C.c_int32()
    lib().orc_g2_table(t.ctypes.data, t.shape[1], t.shape[2], t.shape[0], C.byref(g2), C.byref(df))
    return g2.value, df.value


def pchisq(x, df):
    return lib().orc_pchisq(float(x), int(df))


def chisq_crit(df, alpha):
    return lib().orc_chisq_crit(int(df), float(alpha))


def ci_test(data: Data, x, y, z=(), alpha=0.05):
    z = _z(z)
    r = lib().orc_ci_test(data.ref, x, y, z.ctypes.data, len(z), float(alpha))
    return {"g2": r.g2, "p_value": r.p_value, "df": r.df, "indep": bool(r.indep)}


def ci_batch(data: Data, tests: np.ndarray, max_d: int, alpha: float, n_threads: int = 1):
    """tests: int32 [T, 3 + max_d] rows (x, y, d, z...).  Returns structured arrays."""
    t = np.ascontiguousarray(tests, dtype=np.int32)
    out = (OrcResult * len(t))()
    lib().orc_ci_batch(data.ref, t.ctypes.data, len(t), max_d, float(alpha), n_threads,
                       C.cast(out, C.c_void_p))
    arr = np.ctypeslib.as_array(out)
    return {"g2": arr["g2"].copy(), "p_value": arr["p_value"].copy(), "df": arr["df"].copy(),
            "indep": arr["indep"].astype(bool)}


def enumerate_choices(a, b):
    n = lib().orc_choose(a, b)
    out = np.zeros((max(n, 1), max(b, 1)), dtype=np.int32)
    k = lib().orc_enumerate(a, b, out.ctypes.data, n)
    return out[:k, :b]


def unrank(a, b, rank):
    out = np.zeros(max(b, 1), dtype=np.int32)
    lib().orc_unrank(a, b, int(rank), out.ctypes.data)
    return out[:b]


def pc_skeleton(data: Data, alpha: float, depth: int, n_threads: int = 1):
    n = data.s.n_vars
    e0 = max(1, n * (n - 1) // 2)
    adj = np.zeros((n, n), dtype=np.uint8)
    sep_xy = np.zeros((e0, 2), dtype=np.int32)
    sep_level = np.zeros(e0, dtype=np.int32)
    sep_off = np.zeros(e0 + 1, dtype=np.int64)
    cap = 4 * e0 + 1024
    while True:
        sep_z = np.zeros(cap, dtype=np.int32)
        st = OrcStats()
        rc = lib().orc_pc_skeleton(data.ref, float(alpha), int(depth), int(n_threads), adj.ctypes.data,
                                   sep_xy.ctypes.data, sep_level.ctypes.data, sep_off.ctypes.data,
                                   sep_z.ctypes.data, cap, C.byref(st))
        if rc == -1:
            cap *= 4
            continue
        if rc != 0:
            raise RuntimeError(f"orc_pc_skeleton failed: {rc}")
        break
    k = st.n_sepsets
    seps = {}
    for i in range(k):
        seps[(int(sep_xy[i, 0]), int(sep_xy[i, 1]))] = (
            int(sep_level[i]), tuple(int(v) for v in sep_z[sep_off[i]:sep_off[i + 1]]))
    L = st.n_levels
    return {
        "adj": adj,
        "sepsets": seps,
        "tests_per_level": [int(v) for v in st.tests_per_level[:L]],
        "edges_after_level": [int(v) for v in st.edges_after_level[:L]],
        "num_ci_test": int(st.num_ci_test),
    }


def level0_sample(data: Data, alpha: float, n_pairs: int, n_threads: int):
    out = np.zeros(max(1, n_pairs), dtype=np.int32)
    k = lib().orc_level0_sample(data.ref, float(alpha), int(n_pairs), int(n_threads), out.ctypes.data)
    return out[:k]


def level_sample(data: Data, alpha: float, d: int, adj_before: np.ndarray, max_edges: int, n_threads: int):
    """Runs level d's CheckEdge on the first max_edges edges of adj_before; returns (tests, removed)."""
    a = np.ascontiguousarray(adj_before, dtype=np.uint8)
    rem = C.c_int64()
    t = lib().orc_level_sample(data.ref, float(alpha), int(d), a.ctypes.data, int(max_edges), int(n_threads),
                               C.byref(rem))
    return int(t), int(rem.value)


def level_sample_ref128(data: Data, alpha: float, d: int, adj_before: np.ndarray, max_edges: int, n_threads: int):
    """level_sample under the reference's 128-edge batch schedule (PCStable.cpp:275-324)."""
    a = np.ascontiguousarray(adj_before, dtype=np.uint8)
    rem = C.c_int64()
    t = lib().orc_level_sample_ref128(data.ref, float(alpha), int(d), a.ctypes.data, int(max_edges),
                                      int(n_threads), C.byref(rem))
    return int(t), int(rem.value)


def parent_lists(parents):
    """CSR form (par_off, flat) of per-node ascending parent lists."""
    off = np.zeros(len(parents) + 1, dtype=np.int32)
    flat = []
    for i, ps in enumerate(parents):
        ps = sorted(int(p) for p in ps)
        flat.extend(ps)
        off[i + 1] = len(flat)
    return off, np.ascontiguousarray(np.asarray(flat if flat else [0], dtype=np.int32))


def mle(data: Data, parents, alpha: int = 1):
    """Maximum-likelihood tables and Laplace-smoothed potentials of nodes 0..len(parents)-1
    (ParameterLearning.cpp:24-82): (count_off, counts, potentials)."""
    off, flat = parent_lists(parents)
    n = len(parents)
    count_off = np.zeros(n + 1, dtype=np.int64)
    total = lib().orc_mle(data.ref, n, off.ctypes.data, flat.ctypes.data, int(alpha), count_off.ctypes.data,
                          None, None)
    counts = np.zeros(max(total, 1), dtype=np.int32)
    pots = np.zeros(max(total, 1), dtype=np.float64)
    lib().orc_mle(data.ref, n, off.ctypes.data, flat.ctypes.data, int(alpha), count_off.ctypes.data,
                  counts.ctypes.data, pots.ctypes.data)
    return count_off, counts[:total], pots[:total]
