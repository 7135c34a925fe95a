"""The seeded synthetic generator: determinism and the declared network shapes."""
import numpy as np

from fastpgm_amd import configs, datagen


def test_deterministic_and_padded():
    w = configs.get(0)
    a, b = w.sample(), w.sample()
    assert a.sha256() == b.sha256()
    assert a.stride % 256 == 0 and a.stride >= a.n_rows
    assert not a.cols[:, a.n_rows:].any()
    assert (a.cols[:, : a.n_rows].max(axis=1) < a.card).all()


def test_network_shapes():
    for k in (0, 2, 4):
        w = configs.get(k)
        net = w.network()
        assert len(net.edges) == w.n_edges
        assert max(len(p) for p in net.parents) <= w.max_in
        lo, hi = w.card_law[1], w.card_law[2]
        assert net.card.min() >= lo and net.card.max() <= hi
        rank = np.empty(net.n_vars, dtype=int)
        rank[net.order] = np.arange(net.n_vars)
        assert all(rank[u] < rank[v] for u, v in net.edges)


def test_workload_1_shares_network_of_0():
    a, b = configs.get(0).network(), configs.get(1).network()
    assert np.array_equal(a.card, b.card) and a.edges == b.edges


def test_bootstrap_sizes():
    w = configs.get(4)
    base = datagen.random_sample(1, 5, 1000, 2, 3)
    reps = list(itertools_islice(w.replicates(base), 5))
    for r in reps:
        assert w.rep_rows[0] <= r.n_rows <= w.rep_rows[1]
        assert (r.card == base.card).all()


def itertools_islice(it, n):
    import itertools
    return itertools.islice(it, n)
