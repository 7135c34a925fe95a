"""Seeded synthetic Bayesian networks and forward-sampled uint8 sample columns.

The reference's own sampler (SampleSetGenerator.cpp:20-45) reseeds from the wall
clock on every sample, so it cannot reproduce a workload; this generator is the
project's own and is fully determined by its seed.

Layout produced: one contiguous uint8 buffer ``cols`` of shape ``[n_vars, stride]``,
column-major (``cols[v, r]`` = state of variable ``v`` in row ``r``), ``stride`` =
``n_rows`` rounded up to 256 bytes, padding bytes zero.  States use the identity
coding ``0 .. card[v]-1``, the layout the reference builds in
``Dataset::dataset_columns`` (Dataset.h:32, Dataset.cpp:848-857) narrowed to uint8.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

ROW_ALIGN = 256


def padded_stride(n_rows: int) -> int:
    return max(ROW_ALIGN, (n_rows + ROW_ALIGN - 1) // ROW_ALIGN * ROW_ALIGN)


@dataclass
class Network:
    n_vars: int
    card: np.ndarray                 # int32 [n]
    parents: list                    # list of int arrays, parents of each variable
    order: np.ndarray                # a topological order
    cpts: list = field(default_factory=list)  # per variable: float64 [n_parent_configs, card]

    @property
    def edges(self):
        out = []
        for v, ps in enumerate(self.parents):
            for p in ps:
                out.append((int(p), int(v)))
        return out


@dataclass
class Sample:
    cols: np.ndarray                 # uint8 [n_vars, stride]
    n_rows: int
    card: np.ndarray                 # int32 [n_vars]

    @property
    def stride(self) -> int:
        return int(self.cols.shape[1])

    @property
    def n_vars(self) -> int:
        return int(self.cols.shape[0])

    def sha256(self) -> str:
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.cols[:, : self.n_rows]).tobytes())
        h.update(self.card.astype(np.int32).tobytes())
        return h.hexdigest()


def draw_cardinalities(rng: np.random.Generator, n: int, law) -> np.ndarray:
    """``law`` is ("uniform", lo, hi) or ("inv_k", lo, hi) with P(k) proportional to 1/k."""
    kind, lo, hi = law
    ks = np.arange(lo, hi + 1)
    if kind == "uniform":
        p = np.ones(len(ks))
    elif kind == "inv_k":
        p = 1.0 / ks
    else:
        raise ValueError(f"unknown cardinality law {kind!r}")
    p = p / p.sum()
    return rng.choice(ks, size=n, p=p).astype(np.int32)


def random_dag(rng: np.random.Generator, n: int, n_edges: int, max_in: int) -> tuple:
    """Random DAG with exactly ``n_edges`` edges and in-degree <= ``max_in``.

    A random permutation fixes a topological order; ordered pairs (earlier, later)
    are drawn without replacement and kept while the target's in-degree allows.
    """
    max_possible = 0
    order = rng.permutation(n)
    rank = np.empty(n, dtype=np.int64)
    rank[order] = np.arange(n)
    indeg = np.zeros(n, dtype=np.int64)
    parents = [[] for _ in range(n)]
    for pos in range(n):
        max_possible += min(pos, max_in)
    if n_edges > max_possible:
        raise ValueError("edge count not reachable under the in-degree bound")
    seen = set()
    added = 0
    while added < n_edges:
        batch = rng.integers(0, n, size=(4 * (n_edges - added) + 16, 2))
        for a, b in batch:
            if a == b:
                continue
            u, v = (a, b) if rank[a] < rank[b] else (b, a)
            if (u, v) in seen or indeg[v] >= max_in:
                continue
            seen.add((u, v))
            parents[v].append(int(u))
            indeg[v] += 1
            added += 1
            if added == n_edges:
                break
    parents = [np.array(sorted(p), dtype=np.int64) for p in parents]
    return parents, order


def random_cpts(rng: np.random.Generator, card: np.ndarray, parents: list, conc: float) -> list:
    cpts = []
    for v in range(len(card)):
        n_cfg = int(np.prod(card[parents[v]])) if len(parents[v]) else 1
        cpts.append(rng.dirichlet(np.full(int(card[v]), conc), size=n_cfg))
    return cpts


def make_network(seed: int, n: int, n_edges: int, max_in: int, card_law, conc: float) -> Network:
    ss = np.random.SeedSequence(seed)
    s_dag, s_cpt, _s_rows, _s_boot = ss.spawn(4)
    rng_dag = np.random.Generator(np.random.PCG64(s_dag))
    card = draw_cardinalities(rng_dag, n, card_law)
    parents, order = random_dag(rng_dag, n, n_edges, max_in)
    rng_cpt = np.random.Generator(np.random.PCG64(s_cpt))
    cpts = random_cpts(rng_cpt, card, parents, conc)
    return Network(n_vars=n, card=card, parents=parents, order=order, cpts=cpts)


_GEN = None


def _gen_lib():
    """Native sampler (csrc/gen.c); built on first use when absent."""
    global _GEN
    if _GEN is None:
        import ctypes as C
        import os
        from . import build as B
        if not os.path.exists(B.GEN_LIB):
            B.build_gen()
        L = C.CDLL(B.GEN_LIB)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.fpgm_gen_forward.restype = None
        L.fpgm_gen_forward.argtypes = [i32, i64, i64, vp, vp, vp, vp, vp, vp, C.c_uint64, vp]
        L.fpgm_gen_gather.restype = None
        L.fpgm_gen_gather.argtypes = [i32, i64, vp, i64, vp, vp, i64]
        _GEN = L
    return _GEN


def forward_sample(net: Network, n_rows: int, seed: int, stream: int = 2) -> Sample:
    """Ancestral sampling of ``n_rows`` rows.  Row r's draw for variable v is a
    counter-based uniform of (seed stream, v, r), so the sample is independent of
    the thread count and of how rows are split."""
    ss = np.random.SeedSequence(seed)
    row_seed = int(ss.spawn(4)[stream].generate_state(1, dtype=np.uint64)[0])
    stride = padded_stride(n_rows)
    cols = np.zeros((net.n_vars, stride), dtype=np.uint8)
    order = np.ascontiguousarray(net.order, dtype=np.int32)
    card = np.ascontiguousarray(net.card, dtype=np.int32)
    par_off = np.zeros(net.n_vars + 1, dtype=np.int64)
    for v in range(net.n_vars):
        par_off[v + 1] = par_off[v] + len(net.parents[v])
    par = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.int32) for p in net.parents])
                               if par_off[-1] else np.zeros(1, np.int32), dtype=np.int32)
    cums = [np.cumsum(c, axis=1).ravel() for c in net.cpts]
    cpt_off = np.zeros(net.n_vars, dtype=np.int64)
    acc = 0
    for v in range(net.n_vars):
        cpt_off[v] = acc
        acc += len(cums[v])
    cum = np.ascontiguousarray(np.concatenate(cums), dtype=np.float64)
    if n_rows > 0:
        _gen_lib().fpgm_gen_forward(net.n_vars, n_rows, stride, order.ctypes.data, card.ctypes.data,
                                    par_off.ctypes.data, par.ctypes.data, cpt_off.ctypes.data, cum.ctypes.data,
                                    row_seed, cols.ctypes.data)
    return Sample(cols=cols, n_rows=n_rows, card=net.card.astype(np.int32).copy())


def bootstrap(base: Sample, n_rows: int, rng: np.random.Generator) -> Sample:
    """Rows drawn with replacement from ``base``."""
    idx = np.ascontiguousarray(rng.integers(0, base.n_rows, size=n_rows), dtype=np.int64)
    stride = padded_stride(n_rows)
    cols = np.zeros((base.n_vars, stride), dtype=np.uint8)
    if n_rows > 0:
        _gen_lib().fpgm_gen_gather(base.n_vars, n_rows, base.cols.ctypes.data, base.stride, idx.ctypes.data,
                                   cols.ctypes.data, stride)
    return Sample(cols=cols, n_rows=n_rows, card=base.card.copy())


def sample_from_rows(rows: np.ndarray, card=None) -> Sample:
    """Pack a row-major ``[n_rows, n_vars]`` integer array into the column layout."""
    rows = np.asarray(rows)
    n_rows, n_vars = rows.shape
    if card is None:
        card = rows.max(axis=0).astype(np.int32) + 1 if n_rows else np.ones(n_vars, np.int32)
    stride = padded_stride(n_rows)
    cols = np.zeros((n_vars, stride), dtype=np.uint8)
    cols[:, :n_rows] = rows.T.astype(np.uint8)
    return Sample(cols=cols, n_rows=n_rows, card=np.asarray(card, dtype=np.int32))


def random_sample(seed: int, n_vars: int, n_rows: int, card_lo=2, card_hi=4) -> Sample:
    """Independent uniform columns (no network) for kernel-level parity cases."""
    rng = np.random.default_rng(seed)
    card = rng.integers(card_lo, card_hi + 1, size=n_vars).astype(np.int32)
    stride = padded_stride(n_rows)
    cols = np.zeros((n_vars, stride), dtype=np.uint8)
    for v in range(n_vars):
        cols[v, :n_rows] = rng.integers(0, card[v], size=n_rows)
    return Sample(cols=cols, n_rows=n_rows, card=card)
