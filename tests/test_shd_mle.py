"""Structure comparison (SHD) and maximum-likelihood parameters, without a GPU.

SHD: the host mirror's BNSLComparison / CustomNetwork / FindCompelled (libfpgm_host.so,
through its C entry points) against the values FastPGM's own compiled BNSLComparison
produced (tests/golden/ref_shd.json: the reference's ALARM CPDAG against alarm.bif, and
random learned graphs against random DAGs), and Chickering's labels against the
definition of a compelled edge (same orientation in every DAG of the equivalence class),
enumerated by brute force on small DAGs.

MLE: the oracle's counts against numpy.bincount of an independent index, and its
Laplace-smoothed potentials against the reference's ParameterLearning bit for bit
(tests/golden/ref_mle.npz).
"""
import ctypes as C
import itertools
import json
import os
import subprocess

import numpy as np
import pytest

from fastpgm_amd import datagen
from oracle import oracle as O
from tests.helpers import GOLDEN, alarm_sample

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB = os.path.join(ROOT, "fastpgm_amd", "libfpgm_host.so")
CLI = os.path.join(ROOT, "fastpgm_amd", "fpgm_pcstable")


def host():
    if not os.path.exists(HOST_LIB):
        pytest.fail("fastpgm_amd/libfpgm_host.so is not built (python -m fastpgm_amd.build)")
    L = C.CDLL(HOST_LIB)
    vp, i32, P = C.c_void_p, C.c_int32, C.POINTER(C.c_int32)
    L.fpgm_host_last_error.restype = C.c_char_p
    L.fpgm_host_shd.argtypes = [i32, vp, i32, vp, i32, P]
    L.fpgm_host_shd_bif.argtypes = [C.c_char_p, i32, vp, i32, P]
    L.fpgm_host_bif_arcs.argtypes = [C.c_char_p, P, vp, i32, vp, i32, P]
    L.fpgm_host_compelled.argtypes = [i32, vp, i32, vp]
    for f in ("fpgm_host_shd", "fpgm_host_shd_bif", "fpgm_host_bif_arcs", "fpgm_host_compelled"):
        getattr(L, f).restype = i32
    return L


def shd(L, n, true_arcs, learned):
    ta = np.ascontiguousarray(np.asarray(true_arcs, np.int32).reshape(-1))
    la = np.ascontiguousarray(np.asarray(learned, np.int32).reshape(-1))
    out = C.c_int32()
    assert L.fpgm_host_shd(n, ta.ctypes.data, len(true_arcs), la.ctypes.data, len(learned), C.byref(out)) == 0, \
        L.fpgm_host_last_error()
    return out.value


def golden_shd():
    with open(os.path.join(GOLDEN, "ref_shd.json")) as f:
        return json.load(f)


def test_bif_loader_structure():
    L = host()
    n, na = C.c_int32(), C.c_int32()
    cards = np.zeros(64, np.int32)
    arcs = np.zeros(400, np.int32)
    path = os.path.join(GOLDEN, "alarm.bif").encode()
    assert L.fpgm_host_bif_arcs(path, C.byref(n), cards.ctypes.data, 64, arcs.ctypes.data, 200, C.byref(na)) == 0
    g = golden_shd()["alarm"]
    assert n.value == 37 and na.value == 46
    assert arcs[: 2 * na.value].reshape(-1, 2).tolist() == g["true_arcs"]
    # domain sizes equal the number of values the reference's CSV loader saw in the data
    assert np.array_equal(cards[:37], alarm_sample().card)


def test_shd_alarm_equals_reference():
    L = host()
    g = golden_shd()["alarm"]
    assert g["shd"] == 5
    assert shd(L, g["n"], g["true_arcs"], g["learned"]) == g["shd"]
    la = np.ascontiguousarray(np.asarray(g["learned"], np.int32).reshape(-1))
    out = C.c_int32()
    assert L.fpgm_host_shd_bif(os.path.join(GOLDEN, "alarm.bif").encode(), 37, la.ctypes.data, len(g["learned"]),
                               C.byref(out)) == 0
    assert out.value == g["shd"]


def test_shd_random_graphs_equal_reference():
    L = host()
    cases = golden_shd()["cases"]
    assert len(cases) >= 150
    for c in cases:
        assert shd(L, c["n"], c["true_arcs"], c["learned"]) == c["shd"], c


def test_shd_invariants():
    L = host()
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = int(rng.integers(3, 10))
        order = rng.permutation(n)
        arcs = [[int(order[i]), int(order[j])] for i in range(n) for j in range(i + 1, n) if rng.random() < 0.4]
        if not arcs:
            continue
        # a DAG against itself: 0; the empty pairs never count; one deleted arc: 1
        assert shd(L, n, arcs, [a + [1] for a in arcs]) == 0
        assert shd(L, n, arcs, [a + [1] for a in arcs[1:]] or [[arcs[0][0], arcs[0][1], 1]]) == (len(arcs) > 1)
        # a reversed arc in a DAG-shaped learned graph: 1
        rev = [a + [1] for a in arcs]
        rev[0] = [arcs[0][1], arcs[0][0], 1]
        assert shd(L, n, arcs, rev) == 1


def _v_structures(n, arcs):
    par = [set() for _ in range(n)]
    for a, b in arcs:
        par[b].add(a)
    adj = set()
    for a, b in arcs:
        adj.add((a, b))
        adj.add((b, a))
    out = set()
    for c in range(n):
        for a, b in itertools.combinations(sorted(par[c]), 2):
            if (a, b) not in adj:
                out.add((a, b, c))
    return out


def _acyclic(n, arcs):
    indeg = [0] * n
    ch = [[] for _ in range(n)]
    for a, b in arcs:
        ch[a].append(b)
        indeg[b] += 1
    q = [v for v in range(n) if indeg[v] == 0]
    seen = 0
    while q:
        v = q.pop()
        seen += 1
        for c in ch[v]:
            indeg[c] -= 1
            if indeg[c] == 0:
                q.append(c)
    return seen == n


def test_compelled_labels_match_equivalence_class():
    # Chickering's labels against the definition: an arc is compelled iff every DAG with
    # the same skeleton and v-structures orients it the same way.  In-degree <= 2 keeps
    # FindCompelled off the reference's path that skips labels (a second qualifying parent).
    L = host()
    rng = np.random.default_rng(11)
    checked = 0
    while checked < 60:
        n = int(rng.integers(3, 8))
        order = rng.permutation(n)
        arcs = []
        indeg = [0] * n
        for i in range(n):
            for j in range(i + 1, n):
                c = int(order[j])
                if indeg[c] < 2 and rng.random() < 0.5:
                    arcs.append([int(order[i]), c])
                    indeg[c] += 1
        if not arcs or len(arcs) > 11:
            continue
        vs = _v_structures(n, arcs)
        members = []
        for bits in range(1 << len(arcs)):
            g = [[a, b] if not (bits >> k) & 1 else [b, a] for k, (a, b) in enumerate(arcs)]
            if _acyclic(n, g) and _v_structures(n, g) == vs:
                members.append(g)
        expect = [0 if all(m[k] == arcs[k] for m in members) else 1 for k in range(len(arcs))]
        a = np.ascontiguousarray(np.asarray(arcs, np.int32).reshape(-1))
        labels = np.zeros(len(arcs), np.int32)
        assert L.fpgm_host_compelled(n, a.ctypes.data, len(arcs), labels.ctypes.data) == 0
        assert labels.tolist() == expect, (n, arcs)
        checked += 1


def test_cli_prints_alarm_shd(tmp_path):
    # orientation from the oracle's skeleton on the host, SHD against alarm.bif (no device)
    from tests.test_host import alarm_csv, write_skeleton
    s = alarm_sample()
    g = O.pc_skeleton(O.Data.of(s), 0.05, 1000, n_threads=4)
    sk = tmp_path / "skel"
    write_skeleton(sk, s.n_vars, g)
    csv = alarm_csv(tmp_path)
    r = subprocess.run([CLI, "--skeleton-in", str(sk), "--ref-bif", os.path.join(GOLDEN, "alarm.bif"), str(csv)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "SHD = 5" in r.stdout


def _bincount_tables(cols, card, parents):
    out = []
    for i, ps in enumerate(parents):
        key = cols[i].astype(np.int64)
        size = int(card[i])
        for p in sorted(ps):
            key = key * int(card[p]) + cols[p]
            size *= int(card[p])
        out.append(np.bincount(key, minlength=size))
    return np.concatenate(out)


def test_oracle_mle_against_reference_and_bincount():
    z = np.load(os.path.join(GOLDEN, "ref_mle.npz"), allow_pickle=False)
    s = alarm_sample()
    off, flat = z["alarm_par_off"], z["alarm_par"]
    par = [flat[off[i]:off[i + 1]].tolist() for i in range(s.n_vars)]
    assert sum(len(p) for p in par) == 41  # the arcs of the saved CPDAG
    co, counts, pots = O.mle(O.Data.of(s), par, 1)
    assert np.array_equal(pots, z["alarm_potentials"])  # bit for bit
    assert np.array_equal(counts, _bincount_tables(s.cols[:, : s.n_rows], s.card, par))
    syn = datagen.sample_from_rows(z["syn_cols"].T.astype(np.int32), z["syn_card"])
    off, flat = z["syn_par_off"], z["syn_par"]
    par = [flat[off[i]:off[i + 1]].tolist() for i in range(syn.n_vars)]
    co, counts, pots = O.mle(O.Data.of(syn), par, 1)
    sizes = np.diff(co)
    # every counting path of the device kernel is part of the case: tables with per-wave
    # LDS copies (<= 2048 cells), one LDS copy (<= 8192) and device memory (> 8192)
    assert sizes.min() <= 2048 and ((sizes > 2048) & (sizes <= 8192)).any() and sizes.max() > 8192
    assert np.array_equal(pots, z["syn_potentials"])
    assert np.array_equal(counts, _bincount_tables(syn.cols[:, : syn.n_rows], syn.card, par))
