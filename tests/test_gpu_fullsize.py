"""The workloads at their full stated sizes against the oracle (needs an MI355X).

The 300-variable workload (5 * 10^5 rows) runs the oracle live.  The 1000-variable
workload (10^6 rows, the bench workload) and the 512-replicate bootstrap ensemble take
the oracle minutes, so they are compared with the oracle's results committed by
tests/golden/make_stage_fixtures.py, keyed by the SHA-256 of the data: regenerated data
that differs from the fixture's fails the test.
"""
import numpy as np
import pytest

from fastpgm_amd import _lib, configs
from oracle import oracle as O

from helpers import ensemble_sha256, sepset_digest, stage_fixture

pytestmark = pytest.mark.gpu


def _assert_same(g, o):
    assert np.array_equal(g["adj"], o["adj"])
    assert g["sepsets"] == o["sepsets"]
    assert g["tests_per_level"] == o["tests_per_level"]
    assert g["edges_after_level"] == o["edges_after_level"]
    assert g["num_ci_test"] == o["num_ci_test"]


def _pair_bits(adj):
    iu = np.triu_indices(adj.shape[0], 1)
    return np.packbits(adj[iu].astype(np.uint8))


def _check_stage_fixture(g, f, n):
    assert g["tests_per_level"] == f["tests_per_level"].tolist()
    assert g["edges_after_level"] == f["edges_after_level"].tolist()
    assert g["num_ci_test"] == int(f["num_ci_test"])
    edges = np.array([(i, j) for i in range(n) for j in range(i + 1, n) if g["adj"][i, j]], dtype=np.int32)
    assert np.array_equal(edges.reshape(-1, 2), f["edges"])
    l0 = np.ones((n, n), dtype=np.uint8)
    recs = []
    dmax = f["seps"].shape[1] - 4
    for (x, y), (level, z) in sorted(g["sepsets"].items()):
        if level == 0:
            l0[x, y] = l0[y, x] = 0
        else:
            recs.append([x, y, level, len(z)] + list(z) + [0] * (dmax - len(z)))
    assert np.array_equal(_pair_bits(l0), f["l0_keep"])
    assert np.array_equal(np.array(recs, dtype=np.int32).reshape(-1, 4 + dmax), f["seps"])
    assert sepset_digest(g["sepsets"]) == str(f["sep_digest"])


def test_dag_300v_500k_rows_full_stage():
    # 300 variables, 500 000 rows, alpha 0.01, depth 5 (|Z| <= 4), against a live oracle run
    w = configs.get(2)
    s = w.sample()
    assert s.n_rows == 500_000
    o = O.pc_skeleton(O.Data.of(s), w.alpha, w.depth, n_threads=16)
    with _lib.Context.of(s) as ctx:
        g = ctx.pc_skeleton(w.alpha, w.depth)
    _assert_same(g, o)


def test_dag_1000v_1m_rows_full_stage():
    # the bench workload itself: 1000 variables, 10^6 rows, alpha 0.05, full depth; the
    # default schedule, then the tuned one the bench times
    w = configs.get(3)
    f = stage_fixture(w.name)
    s = w.sample()
    assert s.n_rows == int(f["n_rows"]) == 1_000_000
    assert s.sha256() == str(f["data_sha256"]), "regenerated data differs from the fixture's"
    with _lib.Context.of(s) as ctx:
        g = ctx.pc_skeleton(float(f["alpha"]), int(f["depth"]))
        _check_stage_fixture(g, f, s.n_vars)
        ctx.autotune(float(f["alpha"]), int(f["depth"]))
        g = ctx.pc_skeleton(float(f["alpha"]), int(f["depth"]))
        _check_stage_fixture(g, f, s.n_vars)


def test_bootstrap_512_replicates_full_ensemble():
    # all 512 bootstrap replicates (2 000 - 50 000 rows) of the 100-variable network at the
    # bench depth (levels 0-2), level-synchronously in one context, against 512 independent
    # oracle runs: the edge-frequency matrix, every replicate's skeleton and test count
    w = configs.get(4)
    f = stage_fixture(f"{w.name}-d3")
    base = w.sample()
    reps = list(w.replicates(base))
    assert len(reps) == 512 == len(f["num_ci_test"])
    assert ensemble_sha256(reps) == str(f["data_sha256"]), "regenerated replicates differ from the fixture's"
    with _lib.Context.of_many(reps) as ctx:
        g = ctx.pc_skeleton_multi(float(f["alpha"]), int(f["depth"]))
    assert np.array_equal(g["num_ci_test_per_instance"], f["num_ci_test"])
    assert np.array_equal(np.stack([_pair_bits(a) for a in g["adj"]]), f["adj_bits"])
    assert np.array_equal(g["edge_freq"], f["edge_freq"])
    assert g["num_ci_test"] == int(f["num_ci_test"].sum())
