"""Shared test helpers: fixtures on disk and an independent brute-force PC-stable."""
from __future__ import annotations

import itertools
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def alarm_sample():
    from fastpgm_amd import datagen
    z = np.load(os.path.join(GOLDEN, "alarm_s5000.npz"), allow_pickle=False)
    cols = z["cols"]
    s = datagen.sample_from_rows(cols.T.astype(np.int32), z["card"])
    return s


def alarm_struct():
    with open(os.path.join(GOLDEN, "alarm_s5000_struct.json")) as f:
        return json.load(f)


def brute_counts(cols, card, x, y, z):
    """Histogram of the mixed-radix index by numpy.bincount (independent of the oracle)."""
    key = np.zeros(cols.shape[1], dtype=np.int64)
    for v in z:
        key = key * int(card[v]) + cols[v].astype(np.int64)
    key = (key * int(card[x]) + cols[x]) * int(card[y]) + cols[y]
    dimz = int(np.prod([card[v] for v in z])) if len(z) else 1
    cells = dimz * int(card[x]) * int(card[y])
    return np.bincount(key, minlength=cells).reshape(dimz, int(card[x]), int(card[y]))


def g2_entropy_form(table):
    """G^2 = 2 [sum n log n - sum n_xz log n_xz - sum n_yz log n_yz + sum n_z log n_z] and the
    adjusted df, computed without the per-cell expected-count expression."""
    t = table.astype(np.float64)

    def xlogx(a):
        a = a[a > 0]
        return float((a * np.log(a)).sum())

    nxz = t.sum(axis=2)
    nyz = t.sum(axis=1)
    nz = t.sum(axis=(1, 2))
    g2 = 2.0 * (xlogx(t) - xlogx(nxz) - xlogx(nyz) + xlogx(nz))
    alx = np.maximum((nxz > 0).sum(axis=1), 1)
    aly = np.maximum((nyz > 0).sum(axis=1), 1)
    df = int(((alx - 1) * (aly - 1)).sum())
    return g2, df


def brute_pc_stable(cols, n_rows, card, alpha, depth, test_fn):
    """PC-stable step 1 in plain Python, from the published description of the algorithm
    (Colombo & Maathuis 2014, order-independent skeleton) with the reference's choices:
    lexicographic complete-graph edges, frozen adjacency per level, x-side subsets of
    adj(x)\\{y} in lexicographic order before y-side subsets, first independent set wins,
    continue while max degree - 1 > d.  test_fn(x, y, z) -> bool independent."""
    n = len(card)
    adj = {i: set(range(n)) - {i} for i in range(n)}
    sep = {}
    tests = []
    edges = [(i, j) for i in range(n) for j in range(i + 1, n)]
    removed = []
    for (x, y) in edges:
        if test_fn(x, y, ()):
            removed.append((x, y))
            sep[(x, y)] = (0, ())
    tests.append(len(edges))
    for (x, y) in removed:
        adj[x].discard(y)
        adj[y].discard(x)
    edges = [e for e in edges if e not in set(removed)]
    for d in range(1, depth):
        frozen = {i: sorted(adj[i]) for i in range(n)}
        cnt = 0
        removed = []
        for (x, y) in edges:
            done = False
            for u, other in ((x, y), (y, x)):
                s = [v for v in frozen[u] if v != other]
                if len(s) < d:
                    continue
                for zz in itertools.combinations(s, d):
                    cnt += 1
                    if test_fn(x, y, zz):
                        removed.append((x, y))
                        sep[(x, y)] = (d, tuple(zz))
                        done = True
                        break
                if done:
                    break
        tests.append(cnt)
        for (x, y) in removed:
            adj[x].discard(y)
            adj[y].discard(x)
        edges = [e for e in edges if e not in set(removed)]
        if not (max(len(adj[i]) for i in range(n)) - 1 > d):
            break
    a = np.zeros((n, n), dtype=np.uint8)
    for i in range(n):
        for j in adj[i]:
            a[i, j] = 1
    return a, sep, tests


def sepset_digest(seps):
    """SHA-256 of the sepset map in a canonical text form: one line "x y level z1,z2,..."
    per record, records in ascending (x, y) order."""
    import hashlib
    h = hashlib.sha256()
    for (x, y) in sorted(seps):
        level, z = seps[(x, y)]
        h.update(f"{x} {y} {level} {','.join(str(v) for v in z)}\n".encode())
    return h.hexdigest()


def stage_fixture(name):
    """A committed oracle stage (tests/golden/make_stage_fixtures.py) as a dict of arrays."""
    path = os.path.join(GOLDEN, f"stage_{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def ensemble_sha256(samples):
    """SHA-256 over the valid bytes of every replicate, in order, and the cardinalities."""
    import hashlib
    h = hashlib.sha256()
    for s in samples:
        h.update(np.int64(s.n_rows).tobytes())
        h.update(np.ascontiguousarray(s.cols[:, : s.n_rows]).tobytes())
    h.update(samples[0].card.astype(np.int32).tobytes())
    return h.hexdigest()
