"""HIP path against the oracle, through the C ABI (needs an MI355X)."""
import numpy as np
import pytest

from fastpgm_amd import _lib, configs, datagen
from oracle import oracle as O

from helpers import alarm_sample, alarm_struct

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["0", "1"], ids=["uint8", "packed"])
def column_form(request, monkeypatch):
    # every test under both forms of the columns the counting kernels read: uint8, and
    # packed at 2, 4 or 8 bits per value (by default packing follows the instance size)
    monkeypatch.setenv("FPGM_PACK", request.param)
    return request.param


def _ctx(sample):
    return _lib.Context.of(sample)


def _rand_tests(rng, n_vars, n, dmax):
    out = []
    for _ in range(n):
        d = int(rng.integers(0, dmax + 1))
        vs = rng.choice(n_vars, size=d + 2, replace=False)
        x, y = int(min(vs[0], vs[1])), int(max(vs[0], vs[1]))
        out.append((x, y, tuple(sorted(int(v) for v in vs[2:]))))
    return out


@pytest.mark.parametrize("n_rows", [1, 15, 4099, 70001])
def test_counts_bit_exact(n_rows):
    s = datagen.random_sample(20 + n_rows, 10, n_rows, 1, 5)
    D = O.Data.of(s)
    rng = np.random.default_rng(n_rows)
    with _ctx(s) as ctx:
        for (x, y, z) in _rand_tests(rng, s.n_vars, 25, 4):
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z)), (x, y, z)


@pytest.mark.parametrize("n_rows", [15, 4099, 70001])
def test_counts_bit_exact_packed_widths(monkeypatch, n_rows):
    # cardinalities 1..40: columns packed at 2 bits (<= 4 states), 4 bits (<= 16) and kept
    # at 8 bits, mixed within one test; counts equal the oracle's and the unpacked form's
    s = datagen.random_sample(300 + n_rows, 14, n_rows, 1, 40)
    s.card[:4] = [2, 4, 16, 17]  # the width boundaries
    for v in range(4):
        s.cols[v, :n_rows] = np.random.default_rng(v).integers(0, s.card[v], n_rows)
    D = O.Data.of(s)
    rng = np.random.default_rng(n_rows + 1)
    tests = [t for t in _rand_tests(rng, s.n_vars, 60, 3)
             if np.prod([s.card[v] for v in (t[0], t[1]) + t[2]]) <= 1 << 22]
    monkeypatch.setenv("FPGM_PACK", "1")
    with _ctx(s) as ctx:
        got = [ctx.counts(x, y, z) for (x, y, z) in tests]
    monkeypatch.setenv("FPGM_PACK", "0")
    with _ctx(s) as ctx:
        unpacked = [ctx.counts(x, y, z) for (x, y, z) in tests]
    for (x, y, z), g, u in zip(tests, got, unpacked):
        o = O.counts(D, x, y, z)
        assert np.array_equal(g, o), (x, y, z)
        assert np.array_equal(u, o), (x, y, z)


def test_packed_bytes_accounting(monkeypatch):
    # bytes of the packed columns the counting launches read: below the uint8 figure when
    # columns are packed, equal to it when every column keeps 8 bits
    w = configs.get(2)
    s = w.sample(8_000)
    monkeypatch.setenv("FPGM_PACK", "1")
    with _ctx(s) as ctx:
        ctx.pc_skeleton(w.alpha, 3)
        a = ctx.kernel_stats()
    monkeypatch.setenv("FPGM_PACK", "0")
    with _ctx(s) as ctx:
        ctx.pc_skeleton(w.alpha, 3)
        b = ctx.kernel_stats()
    assert 0 < a["packed_bytes"] < a["alg_bytes"]
    assert a["alg_bytes"] == b["alg_bytes"] and b["packed_bytes"] == b["alg_bytes"]


def test_device_columns_with_dirty_padding():
    # columns handed over in device memory: the padding bytes past n_rows hold 0xFF, which
    # the packed copy must neither count nor let spill into valid rows
    # (device memory from the HIP runtime the library uses; torch carries its own runtime)
    import ctypes as C
    _lib.load()
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    s = datagen.random_sample(41, 9, 5003, 2, 12)
    cols = np.ascontiguousarray(s.cols.copy())
    cols[:, s.n_rows:] = 0xFF
    D = O.Data.of(s)
    rng = np.random.default_rng(3)
    o = O.pc_skeleton(D, 0.05, 1000, n_threads=8)
    dptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(dptr), cols.nbytes) == 0
    try:
        assert hip.hipMemcpy(dptr, cols.ctypes.data, cols.nbytes, 1) == 0  # host to device
        ctx = _lib.Context(n_rows=s.n_rows, card=s.card, device_ptr=dptr.value, n_vars=cols.shape[0],
                           stride=cols.shape[1])
        try:
            for (x, y, z) in _rand_tests(rng, s.n_vars, 30, 3):
                assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z)), (x, y, z)
            g = ctx.pc_skeleton(0.05, 1000)
        finally:
            ctx.close()
    finally:
        hip.hipFree(dptr)
    _assert_same(g, o)


def test_lane_wide_instances_and_routing():
    # tables of 97..128 cells run per lane (64 KiB instance) when a launch has no shared-
    # sink table, and join the shared sink otherwise; 65..96 cells always per lane (48 KiB)
    s = datagen.random_sample(61, 12, 30_011, 4, 8)
    s.card[:6] = [8, 4, 4, 2, 6, 8]
    for v in range(6):
        s.cols[v, :s.n_rows] = np.random.default_rng(100 + v).integers(0, s.card[v], s.n_rows)
    D = O.Data.of(s)
    tests_lw = [t for t in [(0, 1, (2,)), (1, 2, (0,)), (0, 2, (3, 1)), (1, 4, (2,)), (0, 3, (4,))]
                if 64 < np.prod([s.card[v] for v in (t[0], t[1]) + t[2]]) <= 128]
    tests_mixed = tests_lw + [(0, 5, (4,))]  # 384 cells: a shared-sink table in the launch
    assert len(tests_lw) >= 3
    with _ctx(s) as ctx:
        for tests in (tests_lw, tests_mixed):
            r = ctx.ci_batch(tests, alpha=0.05, flags=_lib.FPGM_F_EXACT)
            arr = np.array([[x, y, len(z)] + list(z) + [0] * (2 - len(z)) for (x, y, z) in tests], dtype=np.int32)
            o = O.ci_batch(D, arr, 2, 0.05, n_threads=8)
            assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
            np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)


def _lane_wide_sample(n_rows, seed):
    # cardinalities chosen so that tables of 65..96 cells have c(Y) from 33 to 90: the
    # 48 KiB per-lane sink, whose second wave pair's bias (96) must be added to the key,
    # not ORed into Y.  Variables 4 and 5 depend on variable 0, so that some tables are
    # far from independent.
    s = datagen.random_sample(seed, 8, n_rows, 2, 2)
    s.card[:] = [2, 1, 2, 33, 40, 48, 80, 90]
    rng = np.random.default_rng(seed)
    for v in range(8):
        s.cols[v, :n_rows] = rng.integers(0, s.card[v], n_rows)
    for v in (4, 5):
        keep = rng.random(n_rows) < 0.5
        s.cols[v, :n_rows] = np.where(keep, s.cols[0, :n_rows] * (s.card[v] // 2) + s.cols[v, :n_rows] % 3,
                                      s.cols[v, :n_rows])
    return s


_LANE_WIDE_TESTS = [(0, 3, ()), (0, 4, ()), (0, 5, ()), (1, 6, ()), (1, 7, ()), (2, 5, ()), (2, 4, ()),
                    (0, 4, (1,)), (1, 4, (0,)), (0, 5, (1,)), (1, 3, (2,)), (1, 5, (2,)), (2, 5, (1,))]


@pytest.mark.parametrize("n_rows", [4099, 20_011, 262_143, 300_007])
def test_lane_wide_sink_large_y(n_rows):
    # every test 65..96 cells (no 97..128-cell table in the launch, so the 48 KiB instance
    # counts them below 262 144 rows; from 262 144 rows the shared sink's [cell][64] column
    # layout does), at least 4096 rows so that waves 2 and 3 count rows; through the
    # counts, the batch and a two-instance level 0 on the counting kernel
    s = _lane_wide_sample(n_rows, 17 + n_rows)
    D = O.Data.of(s)
    tests = _LANE_WIDE_TESTS
    for (x, y, z) in tests:
        assert 65 <= int(np.prod([s.card[v] for v in (x, y) + z])) <= 96
    with _ctx(s) as ctx:
        for (x, y, z) in tests:
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z)), (x, y, z)
        r = ctx.ci_batch(tests, alpha=0.05, flags=_lib.FPGM_F_EXACT)
    arr = np.array([[x, y, len(z)] + list(z) + [0] * (1 - len(z)) for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 1, 0.05, n_threads=8)
    assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
    np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)
    assert not o["indep"].all()
    # level 0 of two instances on the counting kernel (no Gram path for this context)
    s2 = _lane_wide_sample(n_rows // 2 + 4096, 5 + n_rows)
    with _lib.Context.of_many([s, s2]) as ctx:
        g = ctx.pc_skeleton_multi(0.05, 1, flags=_lib.FPGM_F_NO_GRAM)
    for i, si in enumerate((s, s2)):
        oi = O.pc_skeleton(O.Data.of(si), 0.05, 1, n_threads=8)
        assert np.array_equal(g["adj"][i], oi["adj"]), i
        assert g["num_ci_test_per_instance"][i] == oi["num_ci_test"]


def test_counts_large_tables_global_path():
    # card 8: d = 3 gives 8^3 * 64 = 32768 cells (beyond the LDS table)
    s = datagen.random_sample(5, 8, 200_003, 8, 8)
    D = O.Data.of(s)
    with _ctx(s) as ctx:
        for z in [(2, 3, 4), (1, 5, 6), (0, 2, 7)]:
            x, y = [v for v in range(8) if v not in z][:2]
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z))


@pytest.mark.parametrize("flags", [_lib.FPGM_F_EXACT, 0])
def test_ci_batch_matches_oracle(flags):
    w = configs.get(1)
    s = w.sample(60_000)
    D = O.Data.of(s)
    rng = np.random.default_rng(4)
    tests = _rand_tests(rng, s.n_vars, 600, 3)
    with _ctx(s) as ctx:
        r = ctx.ci_batch(tests, alpha=0.05, flags=flags)
    arr = np.array([[x, y, len(z)] + list(z) + [0] * (3 - len(z)) for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 3, 0.05, n_threads=8)
    assert np.array_equal(r["df"], o["df"])
    assert np.array_equal(r["indep"], o["indep"])
    if flags & _lib.FPGM_F_EXACT:
        assert r["exact"].all()
        np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)
        np.testing.assert_allclose(r["p_value"], o["p_value"], rtol=1e-9, atol=1e-12)
    else:
        # FP32 statistic within its error bound of the FP64 one
        np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-3, atol=1e-2)


def test_ci_test_single_launch_matches_oracle():
    s = alarm_sample()
    D = O.Data.of(s)
    rng = np.random.default_rng(8)
    with _ctx(s) as ctx:
        for (x, y, z) in _rand_tests(rng, s.n_vars, 20, 3):
            r = ctx.ci_test(x, y, z, alpha=0.05)
            o = O.ci_test(D, x, y, z, alpha=0.05)
            assert r["df"] == o["df"] and r["indep"] == o["indep"]
            assert r["g2"] == pytest.approx(o["g2"], rel=1e-11, abs=1e-9)
            assert r["p_value"] == pytest.approx(o["p_value"], rel=1e-9, abs=1e-12)


def test_ci_group_first_independent():
    # IndependenceResult with a group of conditioning sets (ComputeGSquareXYZGroup,
    # IndependenceTest.cpp:163-287): the index of the first independent set, which
    # PCStable::Testing turns into the sepset; per-test records equal the oracle's
    s = alarm_sample()
    D = O.Data.of(s)
    rng = np.random.default_rng(12)
    firsts = []
    with _ctx(s) as ctx:
        for _ in range(40):
            d, c = int(rng.integers(1, 4)), int(rng.integers(2, 11))
            x, y = sorted(int(v) for v in rng.choice(s.n_vars, 2, replace=False))
            others = [v for v in range(s.n_vars) if v not in (x, y)]
            zs = [tuple(sorted(int(v) for v in rng.choice(others, d, replace=False))) for _ in range(c)]
            res, per = ctx.ci_group(x, y, zs, alpha=0.05)
            arr = np.array([[x, y, d] + list(z) + [0] * (3 - d) for z in zs], dtype=np.int32)
            o = O.ci_batch(D, arr, 3, 0.05, n_threads=8)
            assert np.array_equal(per["indep"], o["indep"]) and np.array_equal(per["df"], o["df"])
            np.testing.assert_allclose(per["g2"], o["g2"], rtol=1e-11, atol=1e-9)
            first = next((m for m in range(c) if o["indep"][m]), -1)
            assert res == {"p_value": -1.0 if first >= 0 else 0.0, "is_independent": first >= 0, "first": first}
            firsts.append(first)
        # both outcomes, and independent sets after the first position, occur
        assert -1 in firsts and max(firsts) >= 1
        # the other dispatches: no Z (first stays -1), one set (first = 0, even when dependent)
        for (x, y, zs) in ((0, 1, [()]), (4, 9, [(2,)]), (12, 30, [(5, 7)])):
            res, per = ctx.ci_group(x, y, zs, alpha=0.05)
            o = O.ci_test(D, x, y, zs[0], alpha=0.05)
            assert res["is_independent"] == o["indep"]
            assert res["first"] == (-1 if not zs[0] else 0)
            assert res["p_value"] == pytest.approx(o["p_value"], rel=1e-9, abs=1e-12)
        with pytest.raises(_lib.FpgmError):
            ctx.ci_group(3, 3, [(1,), (2,)])


def _assert_same(g, o):
    assert np.array_equal(g["adj"], o["adj"])
    assert g["sepsets"] == o["sepsets"]
    assert g["tests_per_level"] == o["tests_per_level"]
    assert g["edges_after_level"] == o["edges_after_level"]
    assert g["num_ci_test"] == o["num_ci_test"]


@pytest.mark.parametrize("mode", [_lib.FPGM_MODE_BATCHED, _lib.FPGM_MODE_PLAIN])
def test_skeleton_alarm(mode):
    s = alarm_sample()
    o = O.pc_skeleton(O.Data.of(s), 0.05, 1000, n_threads=8)
    with _ctx(s) as ctx:
        g = ctx.pc_skeleton(0.05, 1000, mode=mode)
    _assert_same(g, o)
    got = sorted([i, j] for i in range(s.n_vars) for j in range(i + 1, s.n_vars) if g["adj"][i, j])
    assert got == alarm_struct()["adjacencies"]


@pytest.mark.parametrize("key", [0, 1])
def test_skeleton_workloads(key):
    w = configs.get(key)
    s = w.sample()
    o = O.pc_skeleton(O.Data.of(s), w.alpha, w.depth, n_threads=8)
    with _ctx(s) as ctx:
        g = ctx.pc_skeleton(w.alpha, w.depth)
    _assert_same(g, o)


def test_skeleton_round_size_does_not_change_results():
    s = configs.get(0).sample()
    with _ctx(s) as ctx:
        a = ctx.pc_skeleton(0.05, 1000, round_tests=1)
        b = ctx.pc_skeleton(0.05, 1000, round_tests=1 << 20)
        c = ctx.pc_skeleton(0.05, 1000, flags=_lib.FPGM_F_EXACT)
    _assert_same(a, b)
    _assert_same(a, c)
    assert sum(c["in_band_per_level"]) <= sum(c["executed_per_level"])


@pytest.mark.parametrize("key,alpha,depth,rows", [(2, 0.01, 5, 20_000), (3, 0.05, 4, 8_000)])
def test_skeleton_pipelined_rounds_same_results(monkeypatch, key, alpha, depth, rows):
    # FPGM_PIPELINE=1 builds round r+1 while round r runs; the device drops round r+1's
    # tests of edges round r decided (compaction kernel).  Skeleton, sepsets and counted
    # tests must equal the synchronous schedule's and the oracle's; small rounds
    # (round_tests) make many rounds, so that many tests are dropped.
    w = configs.get(key)
    s = w.sample(rows)
    o = O.pc_skeleton(O.Data.of(s), alpha, depth, n_threads=8)
    monkeypatch.setenv("FPGM_PIPELINE", "0")
    with _ctx(s) as ctx:
        a = ctx.pc_skeleton(alpha, depth)
    monkeypatch.setenv("FPGM_PIPELINE", "1")
    with _ctx(s) as ctx:
        b = ctx.pc_skeleton(alpha, depth)
        c = ctx.pc_skeleton(alpha, depth, round_tests=64)
        d = ctx.pc_skeleton(alpha, depth, flags=_lib.FPGM_F_EXACT)
    for g in (a, b, c, d):
        _assert_same(g, o)
    assert sum(c["rounds_per_level"]) > sum(b["rounds_per_level"])


@pytest.mark.parametrize("key,alpha,depth,rows", [(2, 0.01, 5, 20_000), (3, 0.05, 4, 8_000)])
def test_skeleton_concurrent_sink_launches_same_results(monkeypatch, key, alpha, depth, rows):
    # FPGM_CONCURRENT=1 puts a round's slice and device-memory sink launches on streams of
    # their own, joined before the p-value kernel, 0 keeps every launch on one stream; with
    # small rounds too, both schedules must give the oracle's skeleton, sepsets and test
    # counts (the depth-4/5 levels of these samples have slice and device-memory tests)
    w = configs.get(key)
    s = w.sample(rows)
    o = O.pc_skeleton(O.Data.of(s), alpha, depth, n_threads=8)
    out = []
    for conc, slices in (("0", "8"), ("1", "8"), ("1", "1")):
        monkeypatch.setenv("FPGM_CONCURRENT", conc)
        monkeypatch.setenv("FPGM_MAX_SLICES", slices)  # 1: every sliced table to the device sink
        with _ctx(s) as ctx:
            out.append(ctx.pc_skeleton(alpha, depth))
            out.append(ctx.pc_skeleton(alpha, depth, round_tests=256))
    for g in out:
        _assert_same(g, o)


@pytest.mark.parametrize("case", ["dag-1000v-30k", "cards-6-12", "cards-2-8-small"])
def test_grouped_level1_mfma_path(monkeypatch, case):
    # level 1's tests that share the walking endpoint u and the conditioning node z are
    # counted as one-hot int8 products on the MFMA units (FPGM_GROUP=1; off by default); the
    # stage equals the oracle's with the path on and off.  The cases: the 1000-variable
    # network (chunks of 128 columns cut from large groups, row splits), cardinalities 6-12
    # (byte-comparison one-hot form, 36..64 A rows: two fragments), and a small sample whose
    # rows end inside a 32-row step
    if case == "dag-1000v-30k":
        w = configs.get(3)
        s, alpha, depth = w.sample(30_011), w.alpha, 2
    elif case == "cards-6-12":
        net = datagen.make_network(77, 40, 90, 4, ("uniform", 6, 12), 0.4)
        s, alpha, depth = datagen.forward_sample(net, 50_003, 77), 0.05, 3
    else:
        net = datagen.make_network(78, 30, 60, 4, ("uniform", 2, 8), 0.4)
        s, alpha, depth = datagen.forward_sample(net, 4_093, 78), 0.05, 3
    o = O.pc_skeleton(O.Data.of(s), alpha, depth, n_threads=8)
    out = {}
    for on in ("1", "0"):
        monkeypatch.setenv("FPGM_GROUP", on)
        with _ctx(s) as ctx:
            out[on] = ctx.pc_skeleton(alpha, depth)
            out[on]["group"] = ctx.group_stats()
        _assert_same(out[on], o)
    assert out["1"]["group"]["tests"] > 0 and out["0"]["group"]["tests"] == 0
    assert out["1"]["tests_per_level"][1] > 0


@pytest.mark.parametrize("n_rows", [4099, 70_001, 300_007])
def test_bit_plane_tables(monkeypatch, n_rows):
    # |Z| = 1 tables of at most 64 cells (FPGM_BITS_CELLS=64) whose two smallest
    # cardinalities are (2, 2..5), (3, 3..4) or (4, 4) count from bit planes when the columns
    # are packed (planes are built with the packed copy): counts, batch results and a small
    # stage equal the oracle's with the planes on and off.  Every instance of the kernel is
    # reached (C with 1 to 4 planes); 300 007 rows in a small batch are cut into row chunks
    # (chunks start at multiples of 4096 rows, the instance's last word is partial);
    # variable 0 is skewed toward one value
    s = datagen.random_sample(91 + n_rows, 12, n_rows, 2, 2)
    s.card[:] = [2, 2, 3, 3, 4, 4, 5, 6, 8, 10, 2, 4]
    rng = np.random.default_rng(n_rows)
    for v in range(12):
        s.cols[v, :n_rows] = rng.integers(0, s.card[v], n_rows)
    s.cols[0, :n_rows] = np.where(rng.random(n_rows) < 0.9, 1, s.cols[0, :n_rows])
    s.cols[10, :n_rows] = np.where(rng.random(n_rows) < 0.5, s.cols[0, :n_rows], s.cols[10, :n_rows])
    D = O.Data.of(s)
    tests = [(0, 1, (2,)), (0, 2, (1,)), (2, 3, (0,)), (0, 4, (5,)), (4, 10, (2,)), (1, 2, (6,)), (0, 1, (8,)),
             (2, 3, (4,)), (0, 6, (10,)), (1, 7, (11,)), (3, 10, (7,)), (0, 9, (1,)), (4, 5, (0,)), (0, 10, (1,)),
             (2, 4, (5,)), (2, 4, (6,)), (4, 5, (11,)), (0, 6, (7,)), (0, 2, (9,)), (0, 1, (9,)), (3, 5, (4,))]
    for (x, y, z) in tests:
        assert np.prod([s.card[v] for v in (x, y) + z]) <= 64
    arr = np.array([[x, y, 1, z[0]] for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 1, 0.05, n_threads=8)
    ostage = O.pc_skeleton(D, 0.05, 2, n_threads=8)
    monkeypatch.setenv("FPGM_BITS_CELLS", "64")
    for on in ("1", "0"):
        monkeypatch.setenv("FPGM_BITS", on)
        with _ctx(s) as ctx:
            for (x, y, z) in tests:
                assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z)), (on, x, y, z)
            for flags in (_lib.FPGM_F_EXACT, 0):
                r = ctx.ci_batch(tests, alpha=0.05, flags=flags)
                assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
            r = ctx.ci_batch(tests, alpha=0.05, flags=_lib.FPGM_F_EXACT)
            np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)
            _assert_same(ctx.pc_skeleton(0.05, 2), ostage)


@pytest.mark.parametrize("case", ["alarm-5k", "dag-300v-20k", "dag-1000v-8k", "cards-6-12"])
def test_device_rounds_same_results(monkeypatch, case):
    # FPGM_DEVROUNDS=1: the rounds of every level whose tables fit LDS are planned per edge
    # on the host and expanded, routed and reduced on the device; =0: built test by test on
    # the host.  Both equal the oracle's stage: default rounds, small rounds (cut into row
    # chunks with workspace tables), and FP64 decisions.  dag-1000v-8k's deeper levels have
    # tables beyond LDS, so its stage switches to host rounds from level 3 on
    if case == "alarm-5k":
        w = configs.get(0)
        s, alpha, depth = w.sample(), w.alpha, w.depth
    elif case == "dag-300v-20k":
        w = configs.get(2)
        s, alpha, depth = w.sample(20_000), w.alpha, w.depth
    elif case == "dag-1000v-8k":
        w = configs.get(3)
        s, alpha, depth = w.sample(8_000), w.alpha, 5
    else:
        net = datagen.make_network(77, 40, 90, 4, ("uniform", 6, 12), 0.4)
        s, alpha, depth = datagen.forward_sample(net, 50_003, 77), 0.05, 3
    o = O.pc_skeleton(O.Data.of(s), alpha, depth, n_threads=8)
    for on in ("1", "0"):
        monkeypatch.setenv("FPGM_DEVROUNDS", on)
        with _ctx(s) as ctx:
            for rt, fl in ((0, 0), (64, 0), (0, _lib.FPGM_F_EXACT)):
                g = ctx.pc_skeleton(alpha, depth, round_tests=rt, flags=fl)
                _assert_same(g, o)


@pytest.mark.parametrize("on", ["1", "0"])
def test_device_rounds_ensemble(monkeypatch, on):
    # the bootstrap ensemble's level-synchronous stage (levels 0-2) with device and host
    # rounds, default and small rounds: every replicate's skeleton and test count equal
    # its own oracle run
    w, reps = _ensemble(12, max_rows=12_000)
    monkeypatch.setenv("FPGM_DEVROUNDS", on)
    with _lib.Context.of_many(reps) as ctx:
        runs = [ctx.pc_skeleton_multi(w.alpha, 3), ctx.pc_skeleton_multi(w.alpha, 3, round_tests=128)]
    for i, r in enumerate(reps):
        o = O.pc_skeleton(O.Data.of(r), w.alpha, 3, n_threads=8)
        for g in runs:
            assert np.array_equal(g["adj"][i], o["adj"]), i
            assert g["num_ci_test_per_instance"][i] == o["num_ci_test"], i


@pytest.mark.parametrize("key,rows,depth", [(0, None, 1000), (2, 20_000, 5)])
def test_graph_mode_same_results(monkeypatch, key, rows, depth):
    # FPGM_GRAPH=1: every launch sequence's device work captured into a HIP graph (the
    # slot's executable updated in place when the shape repeats) and launched as one; many
    # rounds (small round_tests) replay the graphs with the host's own work between them.
    # Skeleton, sepsets and counts equal the oracle's
    w = configs.get(key)
    s = w.sample(rows)
    o = O.pc_skeleton(O.Data.of(s), w.alpha, depth, n_threads=8)
    monkeypatch.setenv("FPGM_GRAPH", "1")
    with _ctx(s) as ctx:
        for rt in (0, 64, 0):
            _assert_same(ctx.pc_skeleton(w.alpha, depth, round_tests=rt), o)


def test_skeleton_depth_and_alpha():
    w = configs.get(2)
    s = w.sample(20_000)
    D = O.Data.of(s)
    with _ctx(s) as ctx:
        for alpha, depth in ((0.01, 5), (0.05, 2), (0.2, 3)):
            o = O.pc_skeleton(D, alpha, depth, n_threads=8)
            g = ctx.pc_skeleton(alpha, depth)
            _assert_same(g, o)


def test_edge_cases():
    # no rows: every level-0 table is empty, df = 0, all edges removed
    s = datagen.random_sample(1, 5, 0, 2, 3)
    with _ctx(s) as ctx:
        g = ctx.pc_skeleton(0.05, 1000)
    assert g["adj"].sum() == 0 and len(g["sepsets"]) == 10
    # a constant column (cardinality 1)
    rows = np.random.default_rng(0).integers(0, 2, size=(500, 4))
    rows[:, 2] = 0
    s = datagen.sample_from_rows(rows, [2, 2, 1, 2])
    D = O.Data.of(s)
    with _ctx(s) as ctx:
        assert np.array_equal(ctx.counts(0, 1, [2]), O.counts(D, 0, 1, [2]))
        _assert_same(ctx.pc_skeleton(0.05, 1000), O.pc_skeleton(D, 0.05, 1000))


def test_errors_are_status_codes():
    s = datagen.random_sample(2, 4, 100, 2, 3)
    bad = s.cols.copy()
    bad[1, 7] = 9  # value >= cardinality
    with pytest.raises(_lib.FpgmError) as e:
        _lib.Context(bad, s.n_rows, s.card)
    assert e.value.status == _lib.FPGM_EINVAL
    with _ctx(s) as ctx:
        with pytest.raises(_lib.FpgmError) as e:
            ctx.ci_test(0, 1, (), alpha=1.5)
        assert e.value.status == _lib.FPGM_EINVAL
        with pytest.raises(_lib.FpgmError):
            ctx.ci_test(0, 0, ())
        with pytest.raises(_lib.FpgmError):
            ctx.pc_skeleton(-0.1)


def test_plain_sample_keeps_results():
    s = configs.get(1).sample()
    with _ctx(s) as ctx:
        a = ctx.pc_skeleton(0.05, 1000)
        b = ctx.pc_skeleton(0.05, 1000, mode=_lib.FPGM_MODE_PLAIN, plain_sample=50)
    _assert_same(a, b)
    assert b["plain_tests_per_level"][0] == 50
    assert all(p <= 50 for p in b["plain_tests_per_level"])


@pytest.mark.parametrize("alpha", [0.0, 1.0, 0.5])
def test_extreme_alpha(alpha):
    s = configs.get(0).sample()
    D = O.Data.of(s)
    with _ctx(s) as ctx:
        _assert_same(ctx.pc_skeleton(alpha, 3), O.pc_skeleton(D, alpha, 3, n_threads=4))


def test_large_tables_and_high_df():
    # card 8 everywhere: d = 3 tables have 32768 cells (global sink), df up to 3136;
    # d = 4 reaches df > 4096 (critical value beyond the tabulated range)
    s = datagen.random_sample(9, 9, 120_000, 8, 8)
    D = O.Data.of(s)
    rng = np.random.default_rng(3)
    tests = []
    for d in (3, 4):
        for _ in range(6):
            vs = rng.choice(9, size=d + 2, replace=False)
            tests.append((int(min(vs[0], vs[1])), int(max(vs[0], vs[1])), tuple(sorted(int(v) for v in vs[2:]))))
    arr = np.array([[x, y, len(z)] + list(z) + [0] * (4 - len(z)) for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 4, 0.05, n_threads=8)
    with _ctx(s) as ctx:
        for flags in (0, _lib.FPGM_F_EXACT):
            r = ctx.ci_batch(tests, alpha=0.05, flags=flags)
            assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
    assert (o["df"] > 4096).any()


def _ensemble(n_take, max_rows=None):
    w = configs.get(4)
    base = w.sample()          # 50 000-row base sample of the 100-variable network
    reps = []
    for r in w.replicates(base):
        if max_rows is None or r.n_rows <= max_rows:
            reps.append(r)
        if len(reps) == n_take:
            break
    return w, reps


def _check_ensemble(w, reps, depth):
    with _lib.Context.of_many(reps) as ctx:
        g = ctx.pc_skeleton_multi(w.alpha, depth)
    freq = np.zeros((w.n_vars, w.n_vars), dtype=np.int32)
    total = 0
    for i, r in enumerate(reps):
        o = O.pc_skeleton(O.Data.of(r), w.alpha, depth, n_threads=8)
        assert np.array_equal(g["adj"][i], o["adj"]), i
        assert g["num_ci_test_per_instance"][i] == o["num_ci_test"], i
        freq += o["adj"]
        total += o["num_ci_test"]
    assert np.array_equal(g["edge_freq"], freq)
    assert g["num_ci_test"] == total


def test_bootstrap_ensemble_levels_0_1():
    # the first 24 bootstrap replicates (2 000 - 50 000 rows), levels 0 and 1, each
    # checked against its own oracle run; the edge-frequency matrix is their sum
    w, reps = _ensemble(24)
    _check_ensemble(w, reps, 2)


def test_bootstrap_ensemble_full_depth_small_replicates():
    # full depth on replicates of at most 12 000 rows (larger bootstrap replicates
    # of the 50 000-row base inflate G^2 and need millions of tests from level 2 on)
    w, reps = _ensemble(8, max_rows=12_000)
    _check_ensemble(w, reps, w.depth)


@pytest.mark.parametrize("n_rows", [4099, 33, 100_000])
def test_level0_gram_matches_counting_kernel(n_rows):
    # level 0 on the MFMA Gram path (default) and on the counting kernel give the same
    # stage; both equal the oracle
    w = configs.get(2)
    s = w.sample(n_rows)
    D = O.Data.of(s)
    o = O.pc_skeleton(D, w.alpha, 2, n_threads=8)
    with _ctx(s) as ctx:
        a = ctx.pc_skeleton(w.alpha, 2)
        gs = ctx.gram_stats()
        b = ctx.pc_skeleton(w.alpha, 2, flags=_lib.FPGM_F_NO_GRAM)
        gb = ctx.gram_stats()
    assert gs["launches"] == 1 and gb["launches"] == 0
    _assert_same(a, o)
    _assert_same(b, o)


def test_level0_gram_multi_instance():
    # level 0 of every instance of a multi-instance context in one Gram launch (a C block
    # per instance, items over instance x tile pair x row split); instances of different
    # lengths, one of 33 rows (row splits past its end do nothing); equal to the counting
    # kernel's level 0 and to each instance's oracle run
    w, reps = _ensemble(6)
    reps = reps + [datagen.bootstrap(reps[0], 33, np.random.default_rng(1))]
    with _lib.Context.of_many(reps) as ctx:
        a = ctx.pc_skeleton_multi(w.alpha, 1)
        gs = ctx.gram_stats()
        b = ctx.pc_skeleton_multi(w.alpha, 1, flags=_lib.FPGM_F_NO_GRAM)
        gb = ctx.gram_stats()
    assert gs["launches"] == 1 and gb["launches"] == 0
    assert np.array_equal(a["adj"], b["adj"]) and np.array_equal(a["edge_freq"], b["edge_freq"])
    for i, r in enumerate(reps):
        o = O.pc_skeleton(O.Data.of(r), w.alpha, 1, n_threads=8)
        assert np.array_equal(a["adj"][i], o["adj"]), i
        assert a["num_ci_test_per_instance"][i] == o["num_ci_test"]


@pytest.mark.parametrize("card_hi", [8, 12])
def test_level0_gram_card_paths(card_hi):
    # cardinalities <= 8 form the one-hot operands by byte-table lookup, larger ones by
    # byte comparison; both level-0 results equal the counting kernel's and the oracle's.
    # 5003 rows: a partial last 32-row step in every row split; 75 variables (one-hot width
    # not a multiple of the 128-row tile), variable 7 constant (cardinality 1)
    s = datagen.random_sample(77 + card_hi, 75, 5003, 2, card_hi)
    s.card[7] = 1
    s.cols[7, :5003] = 0
    # a few dependent columns so that level 0 keeps some edges
    s.cols[1, :5003] = s.cols[0, :5003] % s.card[1]
    s.cols[3, :5003] = (s.cols[2, :5003] + s.cols[5, :5003]) % s.card[3]
    D = O.Data.of(s)
    o = O.pc_skeleton(D, 0.05, 1, n_threads=8)
    with _ctx(s) as ctx:
        a = ctx.pc_skeleton(0.05, 1)
        gs = ctx.gram_stats()
        b = ctx.pc_skeleton(0.05, 1, flags=_lib.FPGM_F_NO_GRAM)
    assert gs["launches"] == 1
    _assert_same(a, o)
    _assert_same(b, o)


@pytest.mark.parametrize("lo,hi,skew,n_rows,d",
                         [(65, 128, False, 200_003, 1), (33, 256, False, 300_003, 1), (65, 256, True, 300_003, 1),
                          (65, 256, False, 300_003, 2), (257, 512, False, 300_003, 1), (257, 512, True, 300_003, 1),
                          (257, 512, False, 300_003, 2)],
                         ids=["65-128-per-lane", "33-256-col", "65-256-col-skewed", "65-256-d2", "257-512-paged",
                              "257-512-paged-skewed", "257-512-d2"])
def test_mid_tables_many_steps_per_chunk(lo, hi, skew, n_rows, d):
    # hundreds of thousands of rows and thousands of tests of lo..hi cells: every chunk spans
    # tens of 4096-row steps.  Below 262 144 rows, 65..128-cell tables count in the 16-bit
    # per-lane counters (65..96 cells, and 97..128 when the launch has no shared-sink table).
    # From 262 144 rows, |Z| = 1 tables of 65..512 cells count in the per-lane launch's
    # column layouts ([cell][64] up to 128 cells, [cell][32] up to 256, [cell][16] with Z as
    # the page up to 512), |Z| = 2 ones in the shared sink's (same layouts, its own row
    # loop); skewed: nine rows in ten take value 0 in every column, so most lanes of an
    # instruction count into one cell
    w = configs.get(3)
    net = w.network()
    s = datagen.forward_sample(net, n_rows, w.seed)
    card = s.card
    rng = np.random.default_rng(11 + d)
    if skew:
        zero = rng.random(s.n_rows) < 0.9
        s.cols[:, :s.n_rows][:, zero] = 0
    tests = []
    while len(tests) < 3000:
        vs = rng.choice(s.n_vars, size=d + 2, replace=False)
        if lo <= int(np.prod(card[vs])) <= hi:
            x, y = sorted(int(v) for v in vs[:2])
            tests.append((x, y, tuple(sorted(int(v) for v in vs[2:]))))
    D = O.Data.of(s)
    arr = np.array([[x, y, d] + list(z) for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, d, 0.05, n_threads=8)
    with _ctx(s) as ctx:
        r = ctx.ci_batch(tests, alpha=0.05, flags=_lib.FPGM_F_EXACT)
        for (x, y, z) in tests[:4]:
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z))
    assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
    np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("n_rows,n_tests", [(30_001, 400), (300_003, 2000)])
def test_shared_sink_copies_skewed_data(n_rows, n_tests):
    # tables of 257..2000 cells count into LDS histogram copies, lane l in copy l % k: items
    # below 65 536 rows (the first case) use at most 4 copies, longer ones (the second, one
    # chunk per test) up to 8; nine rows in ten take value 0, so most lanes of an
    # instruction hit one cell
    rng = np.random.default_rng(n_rows)
    s = datagen.random_sample(41, 12, n_rows, 6, 12)
    for v in range(s.n_vars):
        col = rng.integers(1, s.card[v], n_rows)
        col[rng.random(n_rows) < 0.9] = 0
        s.cols[v, :n_rows] = col
    tests = []
    while len(tests) < n_tests:
        vs = rng.choice(s.n_vars, size=3, replace=False)
        if 257 <= int(np.prod(s.card[vs])) <= 2000:
            x, y = sorted(int(v) for v in vs[:2])
            tests.append((x, y, (int(vs[2]),)))
    D = O.Data.of(s)
    arr = np.array([[x, y, 1, z[0]] for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 1, 0.05, n_threads=8)
    with _ctx(s) as ctx:
        r = ctx.ci_batch(tests, alpha=0.05, flags=_lib.FPGM_F_EXACT)
        for (x, y, z) in tests[:3]:
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z))
    assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
    np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)


def test_sliced_and_device_memory_sinks():
    # card 8 throughout: |Z| = 2 -> 4096 cells (LDS), 3 -> 32768 (2 z slices),
    # 4 -> 262144 (16 slices: device-memory sink), 5 -> 2^21 (device-memory sink)
    s = datagen.random_sample(21, 9, 50_001, 8, 8)
    D = O.Data.of(s)
    rng = np.random.default_rng(5)
    tests = []
    for d in (2, 3, 4, 5):
        for _ in range(4):
            vs = rng.choice(9, size=d + 2, replace=False)
            tests.append((int(min(vs[0], vs[1])), int(max(vs[0], vs[1])), tuple(sorted(int(v) for v in vs[2:]))))
    arr = np.array([[x, y, len(z)] + list(z) + [0] * (5 - len(z)) for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 5, 0.05, n_threads=8)
    with _ctx(s) as ctx:
        r = ctx.ci_batch(tests, alpha=0.05, flags=_lib.FPGM_F_EXACT)
        for (x, y, z) in tests[4:9]:
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z))
    assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
    np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("mode", ["1", "2"])
@pytest.mark.parametrize("n_rows", [1, 8192, 50_001, 300_007])
def test_partitioned_large_tables(monkeypatch, n_rows, mode):
    # tables beyond LDS counted from rows partitioned by z slice (whole row chunks of 8192
    # rows, a partial one, a single row); card 8 throughout: |Z| = 3 -> 32768 cells (2
    # slices: partitioned in mode 2, z slices in mode 1), 4 -> 262144 (16 slices), 5 -> 2^21
    # (128); |Z| = 2 (4096 cells) stays in LDS
    monkeypatch.setenv("FPGM_PART", mode)
    s = datagen.random_sample(23, 9, n_rows, 8, 8)
    D = O.Data.of(s)
    rng = np.random.default_rng(7)
    tests = []
    for d in (2, 3, 4, 5):
        for _ in range(4):
            vs = rng.choice(9, size=d + 2, replace=False)
            tests.append((int(min(vs[0], vs[1])), int(max(vs[0], vs[1])), tuple(sorted(int(v) for v in vs[2:]))))
    arr = np.array([[x, y, len(z)] + list(z) + [0] * (5 - len(z)) for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 5, 0.05, n_threads=8)
    with _ctx(s) as ctx:
        for flags in (0, _lib.FPGM_F_EXACT):
            r = ctx.ci_batch(tests, alpha=0.05, flags=flags)
            assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
            if flags:
                np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)
        # a parity read keeps the counted table: those tests take the unpartitioned sinks
        for (x, y, z) in tests[4:6]:
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z))


def _latent_sample(seed, n_vars, n_rows, card=8):
    # every column = (shared latent + its own noise in 0..2) mod card: no conditioning set
    # screens a pair off, so every edge survives to the deepest level and the tables grow to
    # card^(|Z| + 2) cells (|Z| = 4: 262 144 cells, 16 z slices)
    rng = np.random.default_rng(seed)
    cols = np.zeros((n_vars, datagen.padded_stride(n_rows)), dtype=np.uint8)
    w = rng.integers(0, card, n_rows)
    for v in range(n_vars):
        cols[v, :n_rows] = (w + rng.integers(0, 3, n_rows)) % card
    return datagen.Sample(cols=cols, n_rows=n_rows, card=np.full(n_vars, card, dtype=np.int32))


@pytest.mark.parametrize("mode", ["0", "1", "2"])
def test_partitioned_tables_in_a_multi_instance_stage(monkeypatch, mode):
    # three instances of different lengths (one row chunk; several; a partial last one) in
    # one level-synchronous stage to depth 5, whose levels 3 and 4 hold 2- and 16-slice
    # tables: z slices and device memory (mode 0), partitioned beyond the slice limit (1),
    # partitioned throughout (2); each instance's skeleton and test count against the oracle
    monkeypatch.setenv("FPGM_PART", mode)
    reps = [_latent_sample(40 + i, 8, n) for i, n in enumerate((6_000, 20_011, 50_001))]
    with _lib.Context.of_many(reps) as ctx:
        g = ctx.pc_skeleton_multi(0.05, 5)
    for i, r in enumerate(reps):
        o = O.pc_skeleton(O.Data.of(r), 0.05, 5, n_threads=8)
        assert o["tests_per_level"][4] > 0
        assert np.array_equal(g["adj"][i], o["adj"]), i
        assert g["num_ci_test_per_instance"][i] == o["num_ci_test"], i


@pytest.mark.parametrize("min_rows", ["0", "1048576"], ids=["chunked", "whole-rows"])
def test_slice_sink_whole_row_items(monkeypatch, min_rows):
    # z-slice tables (8193 .. 131072 cells) with the rows cut into chunks, whose slices are
    # merged in the workspace, and with every work item over all rows: a one-slice table
    # then gets its statistic straight from LDS, a several-slice one merges only slices
    monkeypatch.setenv("FPGM_SLICE_MIN_ROWS", min_rows)
    s = datagen.random_sample(31, 10, 300_001, 5, 8)
    D = O.Data.of(s)
    tests = []
    for x in range(10):
        for y in range(x + 1, 10):
            rest = [v for v in range(10) if v not in (x, y)]
            for d in (2, 3, 4):
                z = tuple(rest[(x + y + k) % len(rest)] for k in range(d))
                if len(set(z)) < d:
                    continue
                z = tuple(sorted(z))
                cells = int(s.card[x]) * int(s.card[y]) * int(np.prod([s.card[v] for v in z]))
                if 8192 < cells <= 131072:
                    tests.append((x, y, z))
    tests = tests[::3]
    n_cells = [int(s.card[x]) * int(s.card[y]) * int(np.prod([s.card[v] for v in z])) for (x, y, z) in tests]
    assert any(c <= 16384 for c in n_cells) and any(c > 16384 for c in n_cells)
    arr = np.array([[x, y, len(z)] + list(z) + [0] * (4 - len(z)) for (x, y, z) in tests], dtype=np.int32)
    o = O.ci_batch(D, arr, 4, 0.05, n_threads=8)
    with _ctx(s) as ctx:
        for flags in (0, _lib.FPGM_F_EXACT):
            r = ctx.ci_batch(tests, alpha=0.05, flags=flags)
            assert np.array_equal(r["df"], o["df"]) and np.array_equal(r["indep"], o["indep"])
            if flags:
                np.testing.assert_allclose(r["g2"], o["g2"], rtol=1e-11, atol=1e-9)
        for (x, y, z) in tests[:2]:
            assert np.array_equal(ctx.counts(x, y, z), O.counts(D, x, y, z))


@pytest.mark.parametrize("key,alpha,depth,rows", [(2, 0.01, 5, 20_000), (3, 0.05, 4, 8_000)])
def test_autotune_changes_time_only(key, alpha, depth, rows):
    # the tuner replays every level under candidate launch shapes and round schedules;
    # the tuned stage, a stage with the saved records and the cleared defaults all equal
    # the oracle
    w = configs.get(key)
    s = w.sample(rows)
    o = O.pc_skeleton(O.Data.of(s), alpha, depth, n_threads=8)
    with _ctx(s) as ctx:
        t = ctx.autotune(alpha, depth)
        assert t["levels"] and t["seconds"] > 0
        for r in t["levels"]:
            assert r["tuned_ms"] <= r["default_ms"] + 1e-9
            assert r["items_per_cu"] > 0 and min(r["wg_per_cu"]) >= 1 and r["pipeline"] in (0, 1)
            assert r["concurrent"] in (0, 1)
        a = ctx.pc_skeleton(alpha, depth)
        ctx.tune_clear()
        b = ctx.pc_skeleton(alpha, depth)
        ctx.tune_set(t["levels"])
        c = ctx.pc_skeleton(alpha, depth)
        with pytest.raises(_lib.FpgmError):
            ctx.tune_set([dict(t["levels"][0], level=99)])
    for g in (a, b, c):
        _assert_same(g, o)
