"""fpgm_ci_batch_noalloc against fpgm_ci_batch through the C ABI (needs an MI355X)."""
import numpy as np
import pytest

from fastpgm_amd import _lib, configs, datagen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tests(rng, card, n, dmax, max_cells=8192):
    out = []
    while len(out) < n:
        d = int(rng.integers(0, dmax + 1))
        vs = rng.choice(len(card), size=d + 2, replace=False)
        if int(np.prod(card[vs])) > max_cells:
            continue
        x, y = sorted(int(v) for v in vs[:2])
        out.append((x, y, tuple(sorted(int(v) for v in vs[2:]))))
    return out


def _same(a, b):
    assert np.array_equal(a["df"], b["df"]) and np.array_equal(a["indep"], b["indep"])
    assert np.array_equal(a["exact"], b["exact"]) and np.array_equal(a["in_band"], b["in_band"])
    np.testing.assert_array_equal(a["g2"], b["g2"])


@pytest.mark.parametrize("pack", ["0", "1"])
def test_noalloc_equals_batch_two_calls(monkeypatch, pack):
    # two calls in a row on one context, the second smaller, each equal to fpgm_ci_batch
    # (bit for bit: same kernels) and to the oracle; every LDS sink (tables up to 8192
    # cells), tables split over several row chunks, and both column forms
    monkeypatch.setenv("FPGM_PACK", pack)
    w = configs.get(3)
    s = datagen.forward_sample(w.network(), 120_011, w.seed)
    D = O.Data.of(s)
    rng = np.random.default_rng(5)
    big = _tests(rng, s.card, 3000, 3)
    small = _tests(rng, s.card, 37, 2)
    with _lib.Context.of(s) as ctx:
        for flags, tests in ((0, big), (_lib.FPGM_F_EXACT, small)):
            a = ctx.ci_batch(tests, alpha=0.05, flags=flags, noalloc=True)
            b = ctx.ci_batch(tests, alpha=0.05, flags=flags)
            _same(a, b)
            dmax = max(len(z) for (_, _, z) in tests)
            arr = np.array([[x, y, len(z)] + list(z) + [0] * (dmax - len(z)) for (x, y, z) in tests], dtype=np.int32)
            o = O.ci_batch(D, arr, dmax, 0.05, n_threads=8)
            assert np.array_equal(a["df"], o["df"]) and np.array_equal(a["indep"], o["indep"])


def test_noalloc_limits_and_alpha_change():
    s = datagen.random_sample(3, 12, 9001, 2, 9)
    s.card[:4] = [9, 9, 9, 9]
    s.cols[:4, :9001] %= 9
    with _lib.Context.of(s) as ctx:
        # 9^4 * 9 * 9 cells > 8192: refused, with a status, before anything runs
        with pytest.raises(_lib.FpgmError) as e:
            ctx.ci_batch([(0, 1, (2, 3, 4, 5))], noalloc=True)
        assert e.value.status == _lib.FPGM_ETOOBIG
        with pytest.raises(_lib.FpgmError) as e:
            ctx.ci_batch([(0, 1, ())] * 4097, noalloc=True)
        assert e.value.status == _lib.FPGM_ETOOBIG
        tests = [(4, 5, (6,)), (0, 7, ()), (1, 2, (3,))]
        for alpha in (0.05, 0.5, 0.05):  # the critical values in the array follow alpha
            _same(ctx.ci_batch(tests, alpha=alpha, noalloc=True), ctx.ci_batch(tests, alpha=alpha))
