"""C++ host mirror (fastpgm_amd/host: Dataset, Network, PCStable) driven through its
command-line driver, without the device: loaders, orientation and the structure file.

The orientation and the structure writer are pinned by the reference's own output: the
CPDAG it saved for ALARM (dataset/alarm/alarm_s5000_struct, committed verbatim) must come
out byte for byte from the oracle's skeleton and separating sets.
"""
import gzip
import os
import subprocess

import numpy as np
import pytest

from fastpgm_amd import configs, datagen
from oracle import oracle as O
from tests.helpers import GOLDEN, alarm_sample

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fastpgm_amd", "fpgm_pcstable")


def run_cli(*args, check=True):
    if not os.path.exists(CLI):
        pytest.fail("fastpgm_amd/fpgm_pcstable is not built (python -m fastpgm_amd.build)")
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError(f"fpgm_pcstable {args} failed ({r.returncode}): {r.stderr}")
    return r


def alarm_csv(tmp_path):
    p = tmp_path / "alarm_s5000"
    with gzip.open(os.path.join(GOLDEN, "alarm_s5000.csv.gz"), "rb") as f:
        p.write_bytes(f.read())
    return p


def write_skeleton(path, n, g):
    with open(path, "w") as f:
        f.write(f"{n}\n")
        for i in range(n):
            for j in range(i + 1, n):
                if g["adj"][i, j]:
                    f.write(f"e {i} {j}\n")
        for (x, y), (_lvl, z) in g["sepsets"].items():
            f.write(f"s {x} {y} {len(z)} " + " ".join(map(str, z)) + "\n")


def read_dump(path):
    raw = open(path, "rb").read()
    nv, nr = np.frombuffer(raw[:8], dtype=np.int32)
    card = np.frombuffer(raw[8: 8 + 4 * nv], dtype=np.int32)
    off = 8 + 4 * nv
    cols = np.frombuffer(raw[off: off + nv * nr], dtype=np.uint8).reshape(nv, nr)
    names = raw[off + nv * nr:].decode().split("\n")[:nv]
    return cols, card, names


def parse_struct(text):
    """(header numbers, undirected pairs, directed arcs) of a structure file."""
    lines = text.rstrip("\n").split("\n")
    und, arcs = [], []
    for ln in lines[1:]:
        body = ln.split(": ", 1)[1]
        if " -- " in body:
            a, b = body.split(" -- ")
            und.append(tuple(sorted((int(a.split("/v")[1]), int(b.split("/v")[1])))))
        else:
            a, b = body.split(" -> ")
            arcs.append((int(a.split("/v")[1]), int(b.split("/v")[1])))
    return lines[0], und, arcs


def is_acyclic(n, arcs):
    indeg = [0] * n
    ch = [[] for _ in range(n)]
    for a, b in arcs:
        ch[a].append(b)
        indeg[b] += 1
    q = [v for v in range(n) if indeg[v] == 0]
    for u in q:
        for c in ch[u]:
            indeg[c] -= 1
            if indeg[c] == 0:
                q.append(c)
    return len(q) == n


def test_cli_built_and_usage():
    r = run_cli("--help")
    assert "usage: fpgm_pcstable" in r.stderr


def test_csv_loader_matches_reference_coding(tmp_path):
    # the host loader codes ALARM exactly as the reference's Dataset::LoadCSVTrainingData
    # compiled from its sources (alarm_s5000.npz was written by it)
    csv = alarm_csv(tmp_path)
    dump = tmp_path / "d.bin"
    run_cli("--load-only", "--dump-data", dump, csv)
    cols, card, names = read_dump(dump)
    ref = np.load(os.path.join(GOLDEN, "alarm_s5000.npz"))
    assert np.array_equal(cols, ref["cols"])
    assert np.array_equal(card, ref["card"])
    assert names == [str(s) for s in ref["names"]]


def test_orientation_and_structure_file_match_reference_alarm(tmp_path):
    csv = alarm_csv(tmp_path)
    s = alarm_sample()
    g = O.pc_skeleton(O.Data.of(s), 0.05, 1000, n_threads=8)
    skel = tmp_path / "skel.txt"
    write_skeleton(skel, s.n_vars, g)
    out = tmp_path / "struct.txt"
    run_cli("--skeleton-in", skel, "-o", out, csv)
    expect = open(os.path.join(GOLDEN, "alarm_s5000_struct.txt")).read()
    assert out.read_text() == expect


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_orientation_invariants_random_networks(tmp_path, seed):
    # random DAG data: the CPDAG keeps the skeleton and its arcs form no cycle; --dag
    # completes it to a DAG over the same skeleton
    n = 14
    net = datagen.make_network(200 + seed, n, 20, 3, ("uniform", 2, 4), 1.0)
    s = datagen.forward_sample(net, 4000, 300 + seed)
    g = O.pc_skeleton(O.Data.of(s), 0.05, 1000, n_threads=4)
    names = ",".join(f"V{i}" for i in range(n))
    rows = s.cols[:, : s.n_rows].T
    csv = tmp_path / "d.csv"
    with open(csv, "w") as f:
        f.write(names + "\n")
        f.write("\n".join(",".join(str(int(v)) for v in r) for r in rows))
    skel = tmp_path / "skel.txt"
    write_skeleton(skel, n, g)
    out = tmp_path / "cpdag.txt"
    run_cli("--skeleton-in", skel, "--int-values", "-o", out, csv)
    head, und, arcs = parse_struct(out.read_text())
    pairs = sorted(set(und) | {tuple(sorted(a)) for a in arcs})
    expect = sorted((i, j) for i in range(n) for j in range(i + 1, n) if g["adj"][i, j])
    assert pairs == expect
    assert head == f"number of nodes = {n}, number of edges = {len(expect)}"
    assert is_acyclic(n, arcs)
    dag = tmp_path / "dag.txt"
    run_cli("--skeleton-in", skel, "--int-values", "--dag", "-o", dag, csv)
    head, und, arcs = parse_struct(dag.read_text())
    if expect:
        assert not und
        assert sorted(tuple(sorted(a)) for a in arcs) == expect
        assert is_acyclic(n, arcs)


def test_add_root_joins_sub_graphs(tmp_path):
    # two disconnected pairs -> ROOT parents one root per part
    csv = tmp_path / "d.csv"
    csv.write_text("A,B,C,D\n0,0,0,0\n1,1,1,1\n")
    skel = tmp_path / "skel.txt"
    skel.write_text("4\ne 0 1\ne 2 3\ns 0 2 0\ns 0 3 0\ns 1 2 0\ns 1 3 0\n")
    out = tmp_path / "s.txt"
    run_cli("--skeleton-in", skel, "--int-values", "--dag", "--add-root", "-o", out, csv)
    text = out.read_text()
    head, und, arcs = parse_struct(text)
    assert head == "number of nodes = 5, number of edges = 4"
    roots = sorted(b for a, b in arcs if a == 4)
    assert len(roots) == 2 and roots[0] in (0, 1) and roots[1] in (2, 3)
    assert "ROOT/v4" in text


def test_libsvm_loader_identity_coding(tmp_path):
    p = tmp_path / "d.libsvm"
    p.write_text("1 1:2 3:1\n0 2:4\n2 1:1 2:1 3:3 \n")
    dump = tmp_path / "d.bin"
    run_cli("--libsvm", "--load-only", "--dump-data", dump, p)
    cols, card, names = read_dump(dump)
    # label at index 0, absent features 0, card = max + 1, names "0".."3"
    assert cols.tolist() == [[1, 0, 2], [2, 0, 1], [0, 4, 1], [1, 0, 3]]
    assert card.tolist() == [3, 3, 5, 4]
    assert names == ["0", "1", "2", "3"]


def test_errors_exit_with_message(tmp_path):
    r = run_cli(tmp_path / "missing.csv", check=False)
    assert r.returncode == 1 and "unable to open file" in r.stderr
    csv = tmp_path / "d.csv"
    csv.write_text("A,A\n0,1\n")
    r = run_cli("--load-only", csv, check=False)
    assert r.returncode == 1 and "duplicate variable names" in r.stderr
    csv.write_text("A,B\n0,1\n")
    r = run_cli("-alpha", "1.5", "--skeleton-in", csv, csv, check=False)
    assert r.returncode == 1 and "alpha" in r.stderr


def test_workload_sample_csv_roundtrip(tmp_path):
    # integer CSV of a synthetic sample loads back to the same columns and cardinalities
    s = configs.get(0).sample(500)
    csv = tmp_path / "w.csv"
    rows = s.cols[:, : s.n_rows].T
    with open(csv, "w") as f:
        f.write(",".join(f"X{i}" for i in range(s.n_vars)) + "\n")
        f.write("\n".join(",".join(str(int(v)) for v in r) for r in rows) + "\n")
    dump = tmp_path / "d.bin"
    run_cli("--int-values", "--load-only", "--dump-data", dump, csv)
    cols, card, _ = read_dump(dump)
    assert np.array_equal(cols, s.cols[:, : s.n_rows])
    assert np.array_equal(card, np.max(s.cols[:, : s.n_rows], axis=1).astype(np.int32) + 1)
