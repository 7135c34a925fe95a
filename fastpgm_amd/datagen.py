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


def forward_sample(net: Network, n_rows: int, seed: int, stream: int = 2) -> Sample:
    """Ancestral sampling of ``n_rows`` rows, one variable at a time over all rows."""
    ss = np.random.SeedSequence(seed)
    rng = np.random.Generator(np.random.PCG64(ss.spawn(4)[stream]))
    stride = padded_stride(n_rows)
    cols = np.zeros((net.n_vars, stride), dtype=np.uint8)
    for v in net.order:
        ps = net.parents[v]
        cfg = np.zeros(n_rows, dtype=np.int64)
        for p in ps:
            cfg = cfg * int(net.card[p]) + cols[p, :n_rows]
        cum = np.cumsum(net.cpts[v], axis=1)
        u = rng.random(n_rows)
        state = (u[:, None] >= cum[cfg]).sum(axis=1)
        np.minimum(state, int(net.card[v]) - 1, out=state)
        cols[v, :n_rows] = state.astype(np.uint8)
    return Sample(cols=cols, n_rows=n_rows, card=net.card.astype(np.int32).copy())


def bootstrap(base: Sample, n_rows: int, rng: np.random.Generator) -> Sample:
    """Rows drawn with replacement from ``base``."""
    idx = rng.integers(0, base.n_rows, size=n_rows)
    stride = padded_stride(n_rows)
    cols = np.zeros((base.n_vars, stride), dtype=np.uint8)
    cols[:, :n_rows] = base.cols[:, idx]
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
