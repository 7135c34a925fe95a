"""ctypes binding of libfpgm_ci.so, the C ABI declared in include/fpgm_ci.h.

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfpgm_ci.so")
_LIB = None

FPGM_OK = 0
FPGM_EINVAL = 1
FPGM_ENOMEM = 2
FPGM_EDEVICE = 3
FPGM_ETOOBIG = 4
FPGM_ETIMEOUT = 5

FPGM_MODE_BATCHED = 0
FPGM_MODE_PLAIN = 1
FPGM_F_EXACT = 1
FPGM_F_FP32ONLY = 2
FPGM_F_NO_GRAM = 4

EXPORTS = [
    "fpgm_ctx_create", "fpgm_ctx_create_device", "fpgm_ctx_destroy", "fpgm_last_error", "fpgm_status_str",
    "fpgm_ci_test", "fpgm_ci_batch", "fpgm_ci_counts", "fpgm_pc_skeleton", "fpgm_sepset_count",
    "fpgm_sepset_get", "fpgm_pchisq", "fpgm_chisq_crit", "fpgm_last_kernel_stats", "fpgm_last_packed_bytes",
    "fpgm_ctx_create_multi", "fpgm_pc_skeleton_multi", "fpgm_last_gram_stats", "fpgm_mle_counts",
    "fpgm_autotune", "fpgm_tune_set", "fpgm_tune_clear", "fpgm_ci_group", "fpgm_ci_batch_noalloc",
    "fpgm_last_group_stats", "fpgm_last_sink_stats",
]


class FpgmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"fpgm status {status}: {msg}")
        self.status = status


class Test(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("d", C.c_int32), ("z_off", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("g2", C.c_double), ("p_value", C.c_double), ("df", C.c_int32), ("indep", C.c_uint8),
                ("in_band", C.c_uint8), ("exact", C.c_uint8), ("pad", C.c_uint8)]


class GroupResult(C.Structure):
    _fields_ = [("p_value", C.c_double), ("is_independent", C.c_uint8), ("pad", C.c_uint8 * 3),
                ("first", C.c_int32)]


class SinkStats(C.Structure):
    _fields_ = [("sink", C.c_int32), ("pad", C.c_int32), ("launches", C.c_int64), ("tests", C.c_int64),
                ("items", C.c_int64), ("ms", C.c_double), ("alg_bytes", C.c_double), ("packed_bytes", C.c_double),
                ("row_tests", C.c_double)]


SINK_NAMES = ["lane", "lane-wide", "shared", "slice", "device", "bits"]


class PcOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("n_threads", C.c_int32), ("group_size", C.c_int32),
                ("verbose", C.c_int32), ("round_tests", C.c_int64), ("max_rounds", C.c_int64),
                ("plain_sample", C.c_int64), ("flags", C.c_uint32), ("band", C.c_float)]


class Stats(C.Structure):
    _fields_ = [("num_ci_test", C.c_int64), ("tests_per_level", C.c_int64 * 64),
                ("edges_after_level", C.c_int64 * 64), ("executed_per_level", C.c_int64 * 64),
                ("in_band_per_level", C.c_int64 * 64), ("rounds_per_level", C.c_int64 * 64),
                ("seconds_per_level", C.c_double * 64), ("plain_tests_per_level", C.c_int64 * 64),
                ("plain_seconds_per_level", C.c_double * 64), ("kernel_seconds_per_level", C.c_double * 64),
                ("n_levels", C.c_int32), ("pad", C.c_int32),
                ("n_sepsets", C.c_int64), ("sep_z_total", C.c_int64), ("ci_seconds", C.c_double),
                ("max_rel_err_g2_fp32", C.c_double)]


class LevelTune(C.Structure):
    _fields_ = [("level", C.c_int32), ("pipeline", C.c_int32), ("concurrent", C.c_int32), ("pad", C.c_int32),
                ("items_per_cu", C.c_int64),
                ("wg_per_cu", C.c_int64 * 5), ("max_slices", C.c_int64), ("round_cap", C.c_int64),
                ("default_ms", C.c_double), ("tuned_ms", C.c_double)]

    def as_dict(self):
        return {"level": self.level, "pipeline": self.pipeline, "concurrent": self.concurrent,
                "items_per_cu": self.items_per_cu,
                "wg_per_cu": list(self.wg_per_cu), "max_slices": self.max_slices, "round_cap": self.round_cap,
                "default_ms": self.default_ms, "tuned_ms": self.tuned_ms}

    @classmethod
    def from_dict(cls, d):
        t = cls()
        t.level, t.pipeline, t.items_per_cu = int(d["level"]), int(d["pipeline"]), int(d["items_per_cu"])
        t.concurrent = int(d.get("concurrent", -1))
        for g in range(5):
            t.wg_per_cu[g] = int(d["wg_per_cu"][g])
        t.max_slices, t.round_cap = int(d["max_slices"]), int(d["round_cap"])
        t.default_ms, t.tuned_ms = float(d.get("default_ms", 0.0)), float(d.get("tuned_ms", 0.0))
        return t


class Sepset(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("level", C.c_int32), ("d", C.c_int32),
                ("z_off", C.c_int64)]


def load(path: str = LIB_PATH):
    """Loads the library (no GPU call is made)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run `python -m fastpgm_amd.build`")
    L = C.CDLL(path)
    vp, i32, i64, f64, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint32
    L.fpgm_ctx_create.argtypes = [vp, i32, i64, i64, vp, C.POINTER(vp)]
    L.fpgm_ctx_create_device.argtypes = [vp, i32, i64, i64, vp, C.POINTER(vp)]
    L.fpgm_ctx_destroy.argtypes = [vp]
    L.fpgm_ctx_destroy.restype = None
    L.fpgm_last_error.argtypes = [vp]
    L.fpgm_last_error.restype = C.c_char_p
    L.fpgm_status_str.argtypes = [i32]
    L.fpgm_status_str.restype = C.c_char_p
    L.fpgm_ci_test.argtypes = [vp, i32, i32, vp, i32, f64, C.POINTER(Result)]
    L.fpgm_ci_batch.argtypes = [vp, vp, i64, vp, i64, f64, u32, vp]
    L.fpgm_ci_counts.argtypes = [vp, i32, i32, vp, i32, vp, i64, C.POINTER(i64)]
    L.fpgm_pc_skeleton.argtypes = [vp, f64, i32, C.POINTER(PcOpts), vp, C.POINTER(Stats)]
    L.fpgm_sepset_count.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.fpgm_sepset_get.argtypes = [vp, vp, i64, vp, i64]
    L.fpgm_pchisq.argtypes = [f64, i32]
    L.fpgm_pchisq.restype = f64
    L.fpgm_chisq_crit.argtypes = [i32, f64]
    L.fpgm_chisq_crit.restype = f64
    L.fpgm_last_kernel_stats.argtypes = [vp, C.POINTER(f64), C.POINTER(f64), C.POINTER(i64)]
    L.fpgm_last_packed_bytes.argtypes = [vp, C.POINTER(f64)]
    L.fpgm_ctx_create_multi.argtypes = [vp, vp, vp, i32, i32, vp, C.POINTER(vp)]
    L.fpgm_last_gram_stats.argtypes = [vp, C.POINTER(f64), C.POINTER(f64), C.POINTER(f64), C.POINTER(i64)]
    L.fpgm_pc_skeleton_multi.argtypes = [vp, f64, i32, C.POINTER(PcOpts), vp, vp, vp, C.POINTER(Stats)]
    L.fpgm_mle_counts.argtypes = [vp, i32, vp, vp, vp, vp, i64]
    L.fpgm_autotune.argtypes = [vp, f64, i32, C.POINTER(PcOpts), vp, i32, C.POINTER(i32), C.POINTER(f64)]
    L.fpgm_tune_set.argtypes = [vp, vp, i32]
    L.fpgm_tune_clear.argtypes = [vp]
    L.fpgm_ci_group.argtypes = [vp, i32, i32, vp, i32, i32, f64, C.POINTER(GroupResult), vp]
    L.fpgm_ci_batch_noalloc.argtypes = [vp, vp, i64, vp, i64, f64, u32, vp]
    L.fpgm_last_group_stats.argtypes = [vp, C.POINTER(f64), C.POINTER(f64), C.POINTER(i64), C.POINTER(i64)]
    L.fpgm_last_sink_stats.argtypes = [vp, vp, i32, C.POINTER(i32)]
    for name in ("fpgm_ctx_create", "fpgm_ctx_create_device", "fpgm_ci_test", "fpgm_ci_batch", "fpgm_ci_counts",
                 "fpgm_pc_skeleton", "fpgm_sepset_count", "fpgm_sepset_get", "fpgm_last_kernel_stats",
                 "fpgm_ctx_create_multi", "fpgm_pc_skeleton_multi", "fpgm_last_gram_stats", "fpgm_mle_counts",
                 "fpgm_autotune", "fpgm_tune_set", "fpgm_tune_clear", "fpgm_last_packed_bytes", "fpgm_ci_group",
                 "fpgm_ci_batch_noalloc", "fpgm_last_group_stats",
                 "fpgm_last_sink_stats"):
        getattr(L, name).restype = i32
    _LIB = L
    return L


class Context:
    """Owns an fpgm_ctx: the uint8 columns resident on the GPU."""

    def __init__(self, cols: np.ndarray | None = None, n_rows: int = 0, card=None, *, device_ptr: int = 0,
                 n_vars: int = 0, stride: int = 0):
        L = load()
        self._L = L
        self.card = np.ascontiguousarray(np.asarray(card, dtype=np.int32))
        h = C.c_void_p()
        if device_ptr:
            st = L.fpgm_ctx_create_device(C.c_void_p(device_ptr), int(n_vars), int(n_rows), int(stride),
                                          self.card.ctypes.data, C.byref(h))
            self.n_vars = int(n_vars)
        else:
            cols = np.ascontiguousarray(cols, dtype=np.uint8)
            st = L.fpgm_ctx_create(cols.ctypes.data, cols.shape[0], int(n_rows), cols.shape[1],
                                   self.card.ctypes.data, C.byref(h))
            self.n_vars = int(cols.shape[0])
        if st != FPGM_OK:
            raise FpgmError(st, L.fpgm_status_str(st).decode())
        self.h = h
        self.n_rows = int(n_rows)
        self.n_inst = 1

    @classmethod
    def of(cls, sample):
        return cls(sample.cols, sample.n_rows, sample.card)

    @classmethod
    def of_many(cls, samples):
        """One context over several samples of the same variables (fpgm_ctx_create_multi)."""
        self = cls.__new__(cls)
        L = load()
        self._L = L
        self.card = np.ascontiguousarray(samples[0].card, dtype=np.int32)
        self._keep = [np.ascontiguousarray(s.cols, dtype=np.uint8) for s in samples]
        ptrs = (C.c_void_p * len(samples))(*[c.ctypes.data for c in self._keep])
        rows = np.array([s.n_rows for s in samples], dtype=np.int64)
        strides = np.array([c.shape[1] for c in self._keep], dtype=np.int64)
        h = C.c_void_p()
        st = L.fpgm_ctx_create_multi(C.cast(ptrs, C.c_void_p), rows.ctypes.data, strides.ctypes.data, len(samples),
                                     self._keep[0].shape[0], self.card.ctypes.data, C.byref(h))
        self._keep = None
        if st != FPGM_OK:
            raise FpgmError(st, L.fpgm_status_str(st).decode())
        self.h = h
        self.n_vars = int(samples[0].n_vars)
        self.n_rows = int(rows.max()) if len(rows) else 0
        self.n_inst = len(samples)
        return self

    def close(self):
        if getattr(self, "h", None):
            self._L.fpgm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != FPGM_OK:
            raise FpgmError(st, self._L.fpgm_last_error(self.h).decode())

    def ci_test(self, x, y, z=(), alpha=0.05):
        z = np.ascontiguousarray(np.asarray(z, dtype=np.int32))
        r = Result()
        self._check(self._L.fpgm_ci_test(self.h, int(x), int(y), z.ctypes.data if len(z) else None, len(z),
                                         float(alpha), C.byref(r)))
        return {"g2": r.g2, "p_value": r.p_value, "df": r.df, "indep": bool(r.indep), "in_band": bool(r.in_band),
                "exact": bool(r.exact)}

    def ci_batch(self, tests, alpha=0.05, flags=0, noalloc=False):
        """tests: sequence of (x, y, z-tuple).  noalloc: through fpgm_ci_batch_noalloc."""
        n = len(tests)
        arr = (Test * max(n, 1))()
        pool = []
        for i, (x, y, z) in enumerate(tests):
            arr[i].x, arr[i].y, arr[i].d, arr[i].z_off = int(x), int(y), len(z), len(pool)
            pool.extend(int(v) for v in z)
        zp = np.ascontiguousarray(np.asarray(pool if pool else [0], dtype=np.int32))
        out = (Result * max(n, 1))()
        fn = self._L.fpgm_ci_batch_noalloc if noalloc else self._L.fpgm_ci_batch
        self._check(fn(self.h, C.cast(arr, C.c_void_p), n, zp.ctypes.data, len(pool), float(alpha), int(flags),
                       C.cast(out, C.c_void_p)))
        a = np.ctypeslib.as_array(out)[:n]
        return {"g2": a["g2"].copy(), "p_value": a["p_value"].copy(), "df": a["df"].copy(),
                "indep": a["indep"].astype(bool), "in_band": a["in_band"].astype(bool),
                "exact": a["exact"].astype(bool)}

    def ci_group(self, x, y, zsets, alpha=0.05):
        """IndependenceResult(x, y, Z, "g square", timer, c_size) with Z the concatenation of
        the c_size sets in zsets (all of one size d; one empty set for d = 0).  Returns the
        group's {"p_value", "is_independent", "first"} and the per-test records."""
        zsets = [tuple(int(v) for v in zs) for zs in zsets]
        d = len(zsets[0]) if zsets else 0
        assert all(len(zs) == d for zs in zsets)
        flat = np.ascontiguousarray(np.asarray([v for zs in zsets for v in zs] or [0], dtype=np.int32))
        n = 1 if d == 0 else len(zsets)
        res = GroupResult()
        out = (Result * max(n, 1))()
        self._check(self._L.fpgm_ci_group(self.h, int(x), int(y), flat.ctypes.data if d else None, d, len(zsets),
                                          float(alpha), C.byref(res), C.cast(out, C.c_void_p)))
        a = np.ctypeslib.as_array(out)[:n]
        return ({"p_value": res.p_value, "is_independent": bool(res.is_independent), "first": res.first},
                {"g2": a["g2"].copy(), "p_value": a["p_value"].copy(), "df": a["df"].copy(),
                 "indep": a["indep"].astype(bool)})

    def counts(self, x, y, z=()):
        z = np.ascontiguousarray(np.asarray(z, dtype=np.int32))
        cells = int(np.prod([self.card[x], self.card[y]] + [self.card[v] for v in z]))
        out = np.zeros(max(cells, 1), dtype=np.int32)
        nc = C.c_int64()
        self._check(self._L.fpgm_ci_counts(self.h, int(x), int(y), z.ctypes.data if len(z) else None, len(z),
                                           out.ctypes.data, len(out), C.byref(nc)))
        dimz = int(np.prod([self.card[v] for v in z])) if len(z) else 1
        return out[:nc.value].reshape(dimz, int(self.card[x]), int(self.card[y]))

    def run_stage(self, alpha, depth=1000, mode=FPGM_MODE_BATCHED, flags=0, round_tests=0, verbose=0,
                  band=0.0, max_rounds=0, plain_sample=0, adj=None):
        """One fpgm_pc_skeleton call; returns (adjacency, raw Stats).  Nothing else is read back."""
        n = self.n_vars
        if adj is None:
            adj = np.zeros((n, n), dtype=np.uint8)
        o = PcOpts(mode=mode, n_threads=1, group_size=1, verbose=verbose, round_tests=round_tests,
                   max_rounds=max_rounds, plain_sample=plain_sample, flags=flags, band=band)
        st = Stats()
        self._check(self._L.fpgm_pc_skeleton(self.h, float(alpha), int(depth), C.byref(o), adj.ctypes.data,
                                             C.byref(st)))
        return adj, st

    def pc_skeleton(self, alpha, depth=1000, mode=FPGM_MODE_BATCHED, flags=0, round_tests=0, verbose=0,
                    band=0.0, max_rounds=0, plain_sample=0):
        adj, st = self.run_stage(alpha, depth, mode, flags, round_tests, verbose, band, max_rounds, plain_sample)
        nr, nz = C.c_int64(), C.c_int64()
        self._check(self._L.fpgm_sepset_count(self.h, C.byref(nr), C.byref(nz)))
        recs = (Sepset * max(1, nr.value))()
        zs = np.zeros(max(1, nz.value), dtype=np.int32)
        self._check(self._L.fpgm_sepset_get(self.h, C.cast(recs, C.c_void_p), nr.value, zs.ctypes.data,
                                            nz.value))
        seps = {}
        for k in range(nr.value):
            r = recs[k]
            seps[(r.x, r.y)] = (r.level, tuple(int(v) for v in zs[r.z_off:r.z_off + r.d]))
        L = st.n_levels
        return {
            "adj": adj,
            "sepsets": seps,
            "num_ci_test": int(st.num_ci_test),
            "tests_per_level": [int(v) for v in st.tests_per_level[:L]],
            "edges_after_level": [int(v) for v in st.edges_after_level[:L]],
            "executed_per_level": [int(v) for v in st.executed_per_level[:L]],
            "in_band_per_level": [int(v) for v in st.in_band_per_level[:L]],
            "rounds_per_level": [int(v) for v in st.rounds_per_level[:L]],
            "seconds_per_level": [float(v) for v in st.seconds_per_level[:L]],
            "plain_tests_per_level": [int(v) for v in st.plain_tests_per_level[:L]],
            "plain_seconds_per_level": [float(v) for v in st.plain_seconds_per_level[:L]],
            "kernel_seconds_per_level": [float(v) for v in st.kernel_seconds_per_level[:L]],
            "ci_seconds": float(st.ci_seconds),
            "max_rel_err_g2_fp32": float(st.max_rel_err_g2_fp32),
        }

    def pc_skeleton_multi(self, alpha, depth=1000, mode=FPGM_MODE_BATCHED, flags=0, round_tests=0, verbose=0,
                          band=0.0, plain_sample=0, want_adj=True):
        """Every instance's CI stage; returns the edge-frequency matrix, per-instance
        skeletons and test counts, and the per-level sums."""
        n, k = self.n_vars, self.n_inst
        freq = np.zeros((n, n), dtype=np.int32)
        adj = np.zeros((k, n, n), dtype=np.uint8) if want_adj else None
        tests = np.zeros(k, dtype=np.int64)
        o = PcOpts(mode=mode, n_threads=1, group_size=1, verbose=verbose, round_tests=round_tests,
                   max_rounds=0, plain_sample=plain_sample, flags=flags, band=band)
        st = Stats()
        self._check(self._L.fpgm_pc_skeleton_multi(self.h, float(alpha), int(depth), C.byref(o), freq.ctypes.data,
                                                   adj.ctypes.data if want_adj else None, tests.ctypes.data,
                                                   C.byref(st)))
        L = st.n_levels
        return {"edge_freq": freq, "adj": adj, "num_ci_test_per_instance": tests,
                "num_ci_test": int(st.num_ci_test), "ci_seconds": float(st.ci_seconds),
                "tests_per_level": [int(v) for v in st.tests_per_level[:L]],
                "executed_per_level": [int(v) for v in st.executed_per_level[:L]],
                "plain_tests_per_level": [int(v) for v in st.plain_tests_per_level[:L]],
                "plain_seconds_per_level": [float(v) for v in st.plain_seconds_per_level[:L]],
                "kernel_seconds_per_level": [float(v) for v in st.kernel_seconds_per_level[:L]]}

    def mle_counts(self, parents):
        """Maximum-likelihood tables N(AB) of nodes 0..len(parents)-1 with the given parent
        lists (sorted here): (count_off int64[n+1], counts int32[total])."""
        n = len(parents)
        off = np.zeros(n + 1, dtype=np.int32)
        flat = []
        for i, ps in enumerate(parents):
            flat.extend(sorted(int(p) for p in ps))
            off[i + 1] = len(flat)
        flat = np.ascontiguousarray(np.asarray(flat if flat else [0], dtype=np.int32))
        count_off = np.zeros(n + 1, dtype=np.int64)
        self._check(self._L.fpgm_mle_counts(self.h, n, off.ctypes.data, flat.ctypes.data, count_off.ctypes.data,
                                            None, 0))
        total = int(count_off[n])
        counts = np.zeros(max(total, 1), dtype=np.int32)
        self._check(self._L.fpgm_mle_counts(self.h, n, off.ctypes.data, flat.ctypes.data, count_off.ctypes.data,
                                            counts.ctypes.data, total))
        return count_off, counts[:total]

    def autotune(self, alpha, depth=1000, flags=0):
        """Per-level launch shapes and round schedule chosen by timing on the device; kept in
        the context for later stages.  Returns {"seconds", "levels": [per-level records]}."""
        o = PcOpts(mode=FPGM_MODE_BATCHED, n_threads=1, group_size=1, flags=flags)
        recs = (LevelTune * 64)()
        n = C.c_int32()
        sec = C.c_double()
        self._check(self._L.fpgm_autotune(self.h, float(alpha), int(depth), C.byref(o), C.cast(recs, C.c_void_p), 64,
                                          C.byref(n), C.byref(sec)))
        return {"seconds": sec.value, "levels": [recs[i].as_dict() for i in range(n.value) if recs[i].level >= 0]}

    def tune_set(self, levels):
        recs = (LevelTune * max(1, len(levels)))(*[LevelTune.from_dict(d) for d in levels])
        self._check(self._L.fpgm_tune_set(self.h, C.cast(recs, C.c_void_p), len(levels)))

    def tune_clear(self):
        self._check(self._L.fpgm_tune_clear(self.h))

    def gram_stats(self):
        ms, ops, pms, n = C.c_double(), C.c_double(), C.c_double(), C.c_int64()
        self._check(self._L.fpgm_last_gram_stats(self.h, C.byref(ms), C.byref(ops), C.byref(pms), C.byref(n)))
        return {"gram_ms": ms.value, "int8_ops": ops.value, "pairs_ms": pms.value, "launches": n.value}

    def group_stats(self):
        ms, ops, t, n = C.c_double(), C.c_double(), C.c_int64(), C.c_int64()
        self._check(self._L.fpgm_last_group_stats(self.h, C.byref(ms), C.byref(ops), C.byref(t), C.byref(n)))
        return {"ms": ms.value, "int8_ops": ops.value, "tests": t.value, "sequences": n.value}

    def sink_stats(self):
        """Per counting sink of the last call: launches, tests, work items, ms, bytes, row-tests."""
        arr = (SinkStats * 8)()
        n = C.c_int32()
        self._check(self._L.fpgm_last_sink_stats(self.h, C.cast(arr, C.c_void_p), 8, C.byref(n)))
        out = {}
        for k in range(min(n.value, 8)):
            r = arr[k]
            out[SINK_NAMES[r.sink]] = {"launches": r.launches, "tests": r.tests, "items": r.items, "ms": r.ms,
                                       "alg_bytes": r.alg_bytes, "packed_bytes": r.packed_bytes,
                                       "row_tests": r.row_tests}
        return out

    def kernel_stats(self):
        ms, b, n = C.c_double(), C.c_double(), C.c_int64()
        self._check(self._L.fpgm_last_kernel_stats(self.h, C.byref(ms), C.byref(b), C.byref(n)))
        pb = C.c_double()
        self._check(self._L.fpgm_last_packed_bytes(self.h, C.byref(pb)))
        return {"kernel_ms": ms.value, "alg_bytes": b.value, "launches": n.value, "packed_bytes": pb.value}


def pchisq(x, df):
    return load().fpgm_pchisq(float(x), int(df))


def chisq_crit(df, alpha):
    return load().fpgm_chisq_crit(int(df), float(alpha))
