"""Pins the CPU restatement (oracle/) against independent computations.

Counts: numpy.bincount of an independently computed index.  G^2/df: the entropy
form of the statistic (no per-cell expected counts).  p-values and critical values:
scipy.stats.chi2.  Enumeration: itertools.combinations.  Skeleton: the reference's
own saved ALARM skeleton and a plain-Python PC-stable on small networks.
"""
import itertools

import numpy as np
import pytest
from scipy import stats

from fastpgm_amd import datagen
from oracle import oracle as O

from helpers import alarm_sample, alarm_struct, brute_counts, brute_pc_stable, g2_entropy_form


@pytest.fixture(scope="module")
def syn():
    return datagen.random_sample(3, 8, 4099, 1, 5)


@pytest.mark.parametrize("d", [0, 1, 2, 3, 4])
def test_counts_match_bincount(syn, d):
    D = O.Data.of(syn)
    rng = np.random.default_rng(d)
    cols = syn.cols[:, : syn.n_rows]
    for _ in range(10):
        vs = rng.choice(syn.n_vars, size=d + 2, replace=False)
        x, y, z = int(vs[0]), int(vs[1]), [int(v) for v in vs[2:]]
        got = O.counts(D, x, y, z)
        exp = brute_counts(cols, syn.card, x, y, z)
        assert np.array_equal(got, exp)
        assert got.sum() == syn.n_rows


def test_counts_first_z_most_significant():
    # 2 rows, Z = (a, b) with cards (2, 3): row values a=1,b=0 -> z index 3; a=0,b=2 -> 2
    rows = np.array([[0, 0, 1, 0], [1, 1, 0, 2]])
    s = datagen.sample_from_rows(rows, [2, 2, 2, 3])
    t = O.counts(O.Data.of(s), 0, 1, [2, 3])
    assert t.shape == (6, 2, 2)
    assert t[3, 0, 0] == 1 and t[2, 1, 1] == 1 and t.sum() == 2


@pytest.mark.parametrize("d", [0, 1, 2, 3])
def test_g2_matches_entropy_form(syn, d):
    D = O.Data.of(syn)
    rng = np.random.default_rng(10 + d)
    for _ in range(10):
        vs = rng.choice(syn.n_vars, size=d + 2, replace=False)
        x, y, z = int(vs[0]), int(vs[1]), [int(v) for v in vs[2:]]
        t = O.counts(D, x, y, z)
        g2, df = O.g2_table(t)
        g2e, dfe = g2_entropy_form(t)
        assert df == dfe
        assert g2 == pytest.approx(g2e, rel=1e-9, abs=1e-9)
        # the statistic is symmetric in X and Y
        g2t, dft = O.g2_table(np.ascontiguousarray(t.transpose(0, 2, 1)))
        assert dft == df and g2t == pytest.approx(g2, rel=1e-12, abs=1e-12)


def test_g2_known_answers():
    # exact product table: G^2 is exactly 0 and df = (2-1)(3-1)
    t = np.outer([1, 3], [2, 2, 4]).reshape(1, 2, 3).astype(np.int32)
    g2, df = O.g2_table(t)
    assert g2 == 0.0 and df == 2
    # 2x2 table [[a, b], [c, d]]: closed form
    a, b, c, d = 30, 10, 5, 55
    t = np.array([[[a, b], [c, d]]], dtype=np.int32)
    n = a + b + c + d
    exp = 2 * sum(o * np.log(o * n / (r * s)) for o, r, s in
                  ((a, a + b, a + c), (b, a + b, b + d), (c, c + d, a + c), (d, c + d, b + d)))
    g2, df = O.g2_table(t)
    assert df == 1 and g2 == pytest.approx(exp, rel=1e-12)
    # an empty z slice contributes nothing; a slice with one x value contributes df 0
    t = np.zeros((3, 2, 2), dtype=np.int32)
    t[1] = [[4, 6], [0, 0]]
    t[2] = [[3, 1], [2, 8]]
    g2, df = O.g2_table(t)
    g2b, dfb = O.g2_table(t[2:3].copy())
    assert df == dfb == 1 and g2 == pytest.approx(g2b, rel=1e-15)


def test_pchisq_matches_scipy():
    for df in [1, 2, 3, 5, 8, 17, 64, 255, 1000, 4096, 60000]:
        for q in [1e-6, 0.01, 0.3, 0.5, 0.9, 0.95, 0.99, 0.999999]:
            x = stats.chi2.ppf(q, df)
            # the prefactor exp(-x + a log x - lgamma(a)) loses ~|a log x| * eps for large a
            rel = 1e-11 if df <= 1000 else 1e-9
            assert O.pchisq(x, df) == pytest.approx(stats.chi2.cdf(x, df), rel=rel, abs=1e-14)
    assert O.pchisq(0.0, 3) == 0.0


@pytest.mark.parametrize("alpha", [0.05, 0.01])
def test_crit_matches_scipy(alpha):
    for df in [1, 2, 7, 33, 500, 4096]:
        assert O.chisq_crit(df, alpha) == pytest.approx(stats.chi2.isf(alpha, df), rel=1e-10)


def test_decision_is_p_greater_than_alpha(syn):
    D = O.Data.of(syn)
    rng = np.random.default_rng(99)
    for _ in range(30):
        vs = rng.choice(syn.n_vars, size=3, replace=False)
        r = O.ci_test(D, int(vs[0]), int(vs[1]), [int(vs[2])], alpha=0.05)
        if r["df"] == 0:
            assert r["indep"] and r["p_value"] == 1.0
            continue
        p = stats.chi2.sf(r["g2"], r["df"])
        assert r["p_value"] == pytest.approx(p, rel=1e-9, abs=1e-13)
        assert r["indep"] == (r["p_value"] > 0.05)


def test_enumeration_is_lexicographic():
    for a in range(1, 9):
        for b in range(1, a + 1):
            got = O.enumerate_choices(a, b)
            exp = np.array(list(itertools.combinations(range(a), b)), dtype=np.int32)
            assert np.array_equal(got, exp)
            for r in range(len(exp)):
                assert np.array_equal(O.unrank(a, b, r), exp[r])


def test_alarm_skeleton_matches_reference_struct_file():
    s = alarm_sample()
    ref = alarm_struct()
    r = O.pc_skeleton(O.Data.of(s), 0.05, 1000, n_threads=4)
    got = sorted([i, j] for i in range(s.n_vars) for j in range(i + 1, s.n_vars) if r["adj"][i, j])
    assert got == ref["adjacencies"]
    assert np.array_equal(r["adj"], r["adj"].T)


def test_skeleton_independent_of_thread_count():
    s = alarm_sample()
    D = O.Data.of(s)
    a = O.pc_skeleton(D, 0.05, 1000, n_threads=1)
    b = O.pc_skeleton(D, 0.05, 1000, n_threads=8)
    assert np.array_equal(a["adj"], b["adj"]) and a["sepsets"] == b["sepsets"]
    assert a["tests_per_level"] == b["tests_per_level"]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_skeleton_matches_plain_python_pc(seed):
    net = datagen.make_network(500 + seed, 9, 11, 3, ("uniform", 2, 3), 1.0)
    s = datagen.forward_sample(net, 3000, 500 + seed)
    cols = s.cols[:, : s.n_rows]
    alpha = 0.05

    def test_fn(x, y, z):
        t = brute_counts(cols, s.card, x, y, list(z))
        g2, df = g2_entropy_form(t)
        if df == 0:
            return True
        return stats.chi2.sf(g2, df) > alpha

    adj, sep, tests = brute_pc_stable(cols, s.n_rows, s.card, alpha, 1000, test_fn)
    r = O.pc_skeleton(O.Data.of(s), alpha, 1000, n_threads=2)
    assert np.array_equal(r["adj"], adj)
    assert r["sepsets"] == sep
    assert r["tests_per_level"] == tests


def test_depth_limits_levels():
    s = alarm_sample()
    r = O.pc_skeleton(O.Data.of(s), 0.05, 2, n_threads=4)
    assert len(r["tests_per_level"]) == 2  # level 0 and level 1 only (d < depth)
    r1 = O.pc_skeleton(O.Data.of(s), 0.05, 1, n_threads=4)
    assert len(r1["tests_per_level"]) == 1
