"""The five PC-stable workloads, by shape.

Each workload is a seeded random DAG, Dirichlet CPTs and forward-sampled uint8
rows.  Master seed ``1000 + k`` for workload ``k``; workload 1 reuses workload 0's
network (same seed) with more rows.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import datagen


@dataclass(frozen=True)
class Workload:
    key: int
    name: str
    n_vars: int
    n_edges: int
    max_in: int
    card_law: tuple
    conc: float
    n_rows: int
    alpha: float
    depth: int
    seed: int
    # bootstrap ensemble (the 512-replicate workload only)
    n_replicates: int = 0
    rep_rows: tuple = (0, 0)

    def network(self) -> datagen.Network:
        return datagen.make_network(self.seed, self.n_vars, self.n_edges, self.max_in,
                                    self.card_law, self.conc)

    def sample(self, n_rows: int | None = None) -> datagen.Sample:
        return datagen.forward_sample(self.network(), n_rows or self.n_rows, self.seed)

    def replicates(self, base: datagen.Sample):
        """Bootstrap replicates: sizes uniform in rep_rows, rows drawn with replacement."""
        ss = np.random.SeedSequence(self.seed)
        rng = np.random.Generator(np.random.PCG64(ss.spawn(4)[3]))
        lo, hi = self.rep_rows
        sizes = rng.integers(lo, hi + 1, size=self.n_replicates)
        for s in sizes:
            yield datagen.bootstrap(base, int(s), rng)


WORKLOADS = {
    0: Workload(0, "alarm-shape-37v-5k", 37, 46, 4, ("uniform", 2, 4), 1.0, 5_000, 0.05, 1000, 1000),
    1: Workload(1, "alarm-shape-37v-200k", 37, 46, 4, ("uniform", 2, 4), 1.0, 200_000, 0.05, 1000, 1000),
    2: Workload(2, "dag-300v-420e-500k", 300, 420, 5, ("uniform", 2, 5), 0.5, 500_000, 0.01, 5, 1002),
    3: Workload(3, "dag-1000v-1500e-1m", 1000, 1500, 6, ("inv_k", 2, 8), 0.3, 1_000_000, 0.05, 1000, 1003),
    4: Workload(4, "bootstrap-512x100v", 100, 130, 4, ("uniform", 2, 6), 1.0, 50_000, 0.05, 1000, 1004,
                n_replicates=512, rep_rows=(2_000, 50_000)),
}


def get(key: int) -> Workload:
    return WORKLOADS[int(key)]
