"""Writes the oracle's CI stages of the full-size workloads as committed fixtures.

The 1000-variable workload (10^6 rows) and the 512-replicate bootstrap ensemble take the
oracle (oracle/liboracle.so, the CPU restatement) minutes to tens of minutes, too long for
a test, so their results are computed once here and stored with the SHA-256 of the data
they came from; the GPU tests regenerate the data, check the SHA-256 and compare.

    make -C oracle && python tests/golden/make_stage_fixtures.py [--threads T] [--procs P]
        [--resume DIR] [3] [4]

Outputs (data only):
  stage_dag-1000v-1500e-1m.npz   data_sha256; alpha, depth; tests_per_level,
      edges_after_level, num_ci_test; l0_keep: the level-0 survivors, one bit per pair of
      the complete graph in lexicographic order (np.packbits); seps: the sepset records of
      levels >= 1 as rows (x, y, level, d, z_1..z_dmax, zero-padded), ascending (x, y);
      edges: the final skeleton's edges (x < y); sep_digest: SHA-256 of every sepset
      record (helpers.sepset_digest), level-0 ones included
  stage_bootstrap-512x100v-d3.npz   data_sha256 (helpers.ensemble_sha256 of the 512
      replicates); alpha, depth (3); edge_freq (100 x 100 int32); num_ci_test per
      replicate; adj_bits: each replicate's skeleton, one bit per pair (np.packbits)
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from fastpgm_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402

from helpers import ensemble_sha256, sepset_digest  # noqa: E402

ENSEMBLE_DEPTH = 3


def pair_bits(adj):
    n = adj.shape[0]
    iu = np.triu_indices(n, 1)
    return np.packbits(adj[iu].astype(np.uint8))


def stage_dag(key, threads):
    w = configs.get(key)
    t0 = time.time()
    s = w.sample()
    print(f"{w.name}: {s.n_vars} x {s.n_rows} generated in {time.time() - t0:.1f} s", flush=True)
    t0 = time.time()
    o = O.pc_skeleton(O.Data.of(s), w.alpha, w.depth, n_threads=threads)
    print(f"{w.name}: oracle stage {o['num_ci_test']} tests in {time.time() - t0:.1f} s", flush=True)
    n = s.n_vars
    # level-0 survivors: the pairs without a level-0 sepset
    l0 = np.ones((n, n), dtype=np.uint8)
    recs = []
    dmax = max([len(z) for (_, z) in o["sepsets"].values()] + [1])
    for (x, y), (level, z) in sorted(o["sepsets"].items()):
        if level == 0:
            l0[x, y] = l0[y, x] = 0
        else:
            recs.append([x, y, level, len(z)] + list(z) + [0] * (dmax - len(z)))
    seps = np.array(recs, dtype=np.int32).reshape(-1, 4 + dmax)
    a = o["adj"]
    edges = np.array([(i, j) for i in range(n) for j in range(i + 1, n) if a[i, j]], dtype=np.int32).reshape(-1, 2)
    out = os.path.join(HERE, f"stage_{w.name}.npz")
    np.savez_compressed(out, data_sha256=np.array(s.sha256()), alpha=np.float64(w.alpha), depth=np.int32(w.depth),
                        n_rows=np.int64(s.n_rows), tests_per_level=np.array(o["tests_per_level"], dtype=np.int64),
                        edges_after_level=np.array(o["edges_after_level"], dtype=np.int64),
                        num_ci_test=np.int64(o["num_ci_test"]), l0_keep=pair_bits(l0), seps=seps, edges=edges,
                        sep_digest=np.array(sepset_digest(o["sepsets"])))
    print(f"wrote {out} ({os.path.getsize(out)} bytes)", flush=True)


def _replicate(args):
    i, r, alpha, threads, resume = args
    path = os.path.join(resume, f"rep{i:04d}.npz") if resume else None
    if path and os.path.exists(path):
        with np.load(path) as f:
            return i, f["adj"], int(f["tests"])
    o = O.pc_skeleton(O.Data.of(r), alpha, ENSEMBLE_DEPTH, n_threads=threads)
    adj = o["adj"].astype(np.uint8)
    if path:
        np.savez(path + ".tmp.npz", adj=adj, tests=np.int64(o["num_ci_test"]))
        os.replace(path + ".tmp.npz", path)
    return i, adj, int(o["num_ci_test"])


def stage_ensemble(threads, procs, resume):
    w = configs.get(4)
    base = w.sample()
    reps = list(w.replicates(base))
    sha = ensemble_sha256(reps)
    n = w.n_vars
    freq = np.zeros((n, n), dtype=np.int32)
    tests = np.zeros(len(reps), dtype=np.int64)
    bits = [None] * len(reps)
    if resume:
        os.makedirs(resume, exist_ok=True)
    t0 = time.time()
    jobs = [(i, r, w.alpha, threads, resume) for i, r in enumerate(reps)]
    done = 0
    if procs > 1:
        import multiprocessing as mp
        pool = mp.get_context("fork").Pool(procs)
        it = pool.imap_unordered(_replicate, jobs)
    else:
        pool, it = None, map(_replicate, jobs)
    for i, adj, nt in it:
        freq += adj
        tests[i] = nt
        bits[i] = pair_bits(adj)
        done += 1
        if done % 32 == 0:
            print(f"{w.name}: {done} replicates, {time.time() - t0:.0f} s", flush=True)
    if pool:
        pool.close()
        pool.join()
    out = os.path.join(HERE, f"stage_{w.name}-d{ENSEMBLE_DEPTH}.npz")
    np.savez_compressed(out, data_sha256=np.array(sha), alpha=np.float64(w.alpha), depth=np.int32(ENSEMBLE_DEPTH),
                        edge_freq=freq, num_ci_test=tests, adj_bits=np.stack(bits))
    print(f"wrote {out} ({os.path.getsize(out)} bytes), {int(tests.sum())} tests", flush=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 8)
    ap.add_argument("--procs", type=int, default=1, help="replicates run in this many processes (ensemble)")
    ap.add_argument("--resume", default="", help="directory keeping each replicate's result (ensemble)")
    ap.add_argument("which", nargs="*", type=int, default=[3, 4])
    a = ap.parse_args()
    for k in a.which:
        if k == 4:
            stage_ensemble(a.threads, a.procs, a.resume)
        else:
            stage_dag(k, a.threads)
