"""Writes the committed golden fixtures from the reference tree and its own compiled code.

Run in the build container only (it reads /root/reference and oracle/_ref):

    make -f oracle/ref.mk && python tests/golden/make_fixtures.py

Outputs (data only: inputs and expected outputs):
  alarm_s5000.npz        ALARM sample (dataset/alarm/alarm_s5000), coded by the reference's
                         own CSV loader (Dataset::LoadCSVTrainingData, first-appearance coding)
  alarm_s5000.csv.gz     the ALARM training file itself (dataset/alarm/alarm_s5000), gzip'd,
                         input of the host loader tests
  alarm_s5000_struct.json  the 44 adjacencies of dataset/alarm/alarm_s5000_struct
  alarm_s5000_struct.txt the reference's saved CPDAG (dataset/alarm/alarm_s5000_struct) verbatim,
                         expected output of orientation + structure file
  ref_choices.json       ChoiceGenerator series for small (a, b)
  ref_counts.npz         Counts2D / Counts3D tables and marginals for sampled tests
  ref_counts_group.npz   Counts3DGroup tables for sampled groups
  alarm.bif              the true ALARM network (dataset/alarm/alarm.bif) verbatim
  ref_shd.json           BNSLComparison::GetSHD of the reference's ALARM CPDAG against
                         alarm.bif, and of random learned graphs against random DAGs
                         (cases on which the reference itself crashes are left out)
  ref_mle.npz            ParameterLearning potentials (Laplace alpha 1) for ALARM with the
                         parents of the saved CPDAG, and for a synthetic sample with random
                         parent sets (tables beyond the LDS size included)
  ref_alarm_param.txt.gz the reference's SaveBNParameter file for that ALARM network
"""
from __future__ import annotations

import ctypes as C
import gzip
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("FPGM_REFERENCE", "/root/reference")
sys.path.insert(0, ROOT)

from fastpgm_amd import datagen  # noqa: E402

import shutil  # noqa: E402
import subprocess  # noqa: E402


def ref_lib():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libfpgm_ref.so"))
    vp = C.c_void_p
    L.ref_csv_load.restype = vp
    L.ref_csv_load.argtypes = [C.c_char_p]
    L.ref_csv_dims.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.ref_csv_fetch.argtypes = [vp, vp, C.c_int64, vp]
    L.ref_csv_free.argtypes = [vp]
    L.ref_counts.argtypes = [vp, C.c_int64, C.c_int32, C.c_int32, vp, C.c_int32, C.c_int32, vp, C.c_int32,
                             vp, vp, vp, vp]
    L.ref_counts_group.argtypes = [vp, C.c_int64, C.c_int32, C.c_int32, vp, C.c_int32, C.c_int32, vp,
                                   C.c_int32, C.c_int32, vp]
    L.ref_choices.restype = C.c_int64
    L.ref_choices.argtypes = [C.c_int32, C.c_int32, vp, C.c_int64]
    L.ref_mle.restype = C.c_int64
    L.ref_mle.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_int32, vp, vp, vp, C.c_int32, vp, C.c_int64,
                          C.c_char_p]
    L.ref_shd.restype = C.c_int32
    L.ref_shd.argtypes = [C.c_int32, vp, C.c_int32, vp, C.c_int32]
    return L


def ref_isolated(fn, *args):
    """Runs fn(*args) (an int-returning reference call) in a forked child: some inputs make
    the reference itself crash (undefined behaviour); those give None."""
    r, w = os.pipe()
    pid = os.fork()
    if pid == 0:
        os.close(r)
        v = fn(*args)
        os.write(w, int(v).to_bytes(4, "little", signed=True))
        os._exit(0)
    os.close(w)
    data = os.read(r, 4)
    os.close(r)
    _, status = os.waitpid(pid, 0)
    if status != 0 or len(data) != 4:
        return None
    return int.from_bytes(data, "little", signed=True)


def struct_records(text):
    """(a, b, directed) records of a structure file, in file order."""
    rec = []
    for ln in text.split("\n")[1:]:
        m = re.findall(r"/v(\d+)", ln)
        if len(m) != 2:
            continue
        a, b = int(m[0]), int(m[1])
        if " -- " in ln:
            rec.append([min(a, b), max(a, b), 0])
        else:
            rec.append([a, b, 1])
    return rec


def bif_arcs(path):
    """(n_nodes, arcs in loader order) of a BIF file: variables in declaration order, one arc
    per listed parent of each probability block (CustomNetwork.cpp:52-160)."""
    names, arcs = [], []
    for ln in open(path):
        t = ln.strip().split(" ")
        if t[0] == "variable":
            names.append(t[1])
        elif t[0] == "probability":
            node = names.index(t[2])
            if t[3] != ")":
                pos = 4
                while t[pos] != ")":
                    arcs.append([names.index(t[pos].rstrip(",")), node])
                    pos += 1
    return len(names), arcs


def csr(parents):
    off = np.zeros(len(parents) + 1, np.int32)
    flat = []
    for i, ps in enumerate(parents):
        flat.extend(sorted(ps))
        off[i + 1] = len(flat)
    return off, np.array(flat if flat else [0], np.int32)


def shd_fixtures(L):
    bif = os.path.join(REF, "dataset", "alarm", "alarm.bif")
    shutil.copyfile(bif, os.path.join(HERE, "alarm.bif"))
    n, arcs = bif_arcs(bif)
    learned = struct_records(open(os.path.join(REF, "dataset", "alarm", "alarm_s5000_struct")).read())
    ta = np.array(arcs, np.int32).ravel()
    la = np.array(learned, np.int32).ravel()
    alarm = {"n": n, "true_arcs": arcs, "learned": learned,
             "shd": ref_isolated(L.ref_shd, n, ta.ctypes.data, len(arcs), la.ctypes.data, len(learned))}
    rng = np.random.default_rng(2024)
    cases = []
    while len(cases) < 160:
        n = int(rng.integers(3, 12))
        order = rng.permutation(n)
        ar = [(int(order[i]), int(order[j])) for i in range(n) for j in range(i + 1, n) if rng.random() < 0.35]
        if not ar:
            continue
        rng.shuffle(ar)
        skel = set(tuple(sorted(a)) for a in ar)
        for _ in range(int(rng.integers(0, 3))):
            a, b = rng.choice(n, 2, replace=False)
            skel.add((int(min(a, b)), int(max(a, b))))
        skel = sorted(e for e in skel if rng.random() < 0.9)
        if not skel:
            continue
        mode = int(rng.integers(0, 3))
        und, dirs = [], []
        for a, b in skel:
            if mode == 0 or (mode == 1 and rng.random() < 0.4):
                und.append([a, b, 0])
            elif rng.random() < 0.5:
                dirs.append([a, b, 1])
            else:
                dirs.append([b, a, 1])
        rec = und + dirs
        ta = np.array(ar, np.int32).ravel()
        la = np.array(rec, np.int32).ravel()
        v = ref_isolated(L.ref_shd, n, ta.ctypes.data, len(ar), la.ctypes.data, len(rec))
        if v is None:
            continue
        cases.append({"n": n, "true_arcs": [list(a) for a in ar], "learned": rec, "shd": v})
    json.dump({"alarm": alarm, "cases": cases}, open(os.path.join(HERE, "ref_shd.json"), "w"))


def mle_fixtures(L, cols, nrow, card):
    csv = os.path.join(REF, "dataset", "alarm", "alarm_s5000")
    learned = struct_records(open(os.path.join(REF, "dataset", "alarm", "alarm_s5000_struct")).read())
    nv = cols.shape[0]
    alarm_par = [[] for _ in range(nv)]
    for a, b, d in learned:
        if d:
            alarm_par[b].append(a)
    off, flat = csr(alarm_par)
    cap = 1 << 20
    pots = np.zeros(cap, np.float64)
    h = L.ref_csv_load(csv.encode())
    param_path = os.path.join(HERE, "_alarm_param.tmp")
    k = L.ref_mle(h, None, 0, 0, 0, None, off.ctypes.data, flat.ctypes.data, 1, pots.ctypes.data, cap,
                  param_path.encode())
    L.ref_csv_free(h)
    alarm_pots = pots[:k].copy()
    with open(param_path, "rb") as f, open(os.path.join(HERE, "ref_alarm_param.txt.gz"), "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as g:
            g.write(f.read())
    os.remove(param_path)
    # synthetic: 14 variables with cardinality 2..6, random parent sets of up to 6 parents
    syn = datagen.random_sample(31, 14, 20011, 2, 6)
    rng = np.random.default_rng(5)
    syn_par = []
    for i in range(syn.n_vars):
        kk = int(rng.integers(0, 7))
        others = [v for v in range(syn.n_vars) if v != i]
        syn_par.append(sorted(int(v) for v in rng.choice(others, kk, replace=False)))
    off2, flat2 = csr(syn_par)
    sc = np.ascontiguousarray(syn.cols)
    k2 = L.ref_mle(None, sc.ctypes.data, sc.shape[1], syn.n_vars, syn.n_rows, syn.card.ctypes.data,
                   off2.ctypes.data, flat2.ctypes.data, 1, pots.ctypes.data, cap, None)
    np.savez_compressed(os.path.join(HERE, "ref_mle.npz"),
                        alarm_par_off=off, alarm_par=flat, alarm_potentials=alarm_pots,
                        syn_cols=syn.cols[:, : syn.n_rows], syn_card=syn.card, syn_par_off=off2, syn_par=flat2,
                        syn_potentials=pots[:k2].copy())



def first_appearance_coding(path):
    """Restatement of Dataset.cpp:459-467 (value ids in order of first appearance per column)."""
    with open(path) as f:
        text = f.read()
    lines = text.split("\n")
    header = lines[0].rstrip().split(",")
    rows = [ln.rstrip().split(",") for ln in lines[1:] if ln.strip() != ""]
    maps = [dict() for _ in header]
    arr = np.zeros((len(rows), len(header)), dtype=np.int32)
    for r, row in enumerate(rows):
        for i, v in enumerate(row):
            m = maps[i]
            if v not in m:
                m[v] = len(m)
            arr[r, i] = m[v]
    card = np.array([len(m) for m in maps], dtype=np.int32)
    return header, arr, card


def main():
    L = ref_lib()
    csv = os.path.join(REF, "dataset", "alarm", "alarm_s5000")
    h = L.ref_csv_load(csv.encode())
    nv, nr = C.c_int32(), C.c_int32()
    L.ref_csv_dims(h, C.byref(nv), C.byref(nr))
    stride = datagen.padded_stride(nr.value)
    cols = np.zeros((nv.value, stride), dtype=np.uint8)
    card = np.zeros(nv.value, dtype=np.int32)
    L.ref_csv_fetch(h, cols.ctypes.data, stride, card.ctypes.data)
    L.ref_csv_free(h)
    names, arr, card_py = first_appearance_coding(csv)
    assert np.array_equal(card, card_py), "CSV coding restatement disagrees with the reference loader (card)"
    assert np.array_equal(cols[:, : nr.value], arr.T.astype(np.uint8)), "CSV coding restatement disagrees"
    np.savez_compressed(os.path.join(HERE, "alarm_s5000.npz"), cols=cols[:, : nr.value], card=card,
                        names=np.array(names))

    with open(csv, "rb") as f, open(os.path.join(HERE, "alarm_s5000.csv.gz"), "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as g:
            g.write(f.read())
    with open(os.path.join(REF, "dataset", "alarm", "alarm_s5000_struct"), "rb") as f:
        struct_text = f.read()
    with open(os.path.join(HERE, "alarm_s5000_struct.txt"), "wb") as f:
        f.write(struct_text)

    edges = []
    with open(os.path.join(REF, "dataset", "alarm", "alarm_s5000_struct")) as f:
        lines = f.read().split("\n")
    for ln in lines[1:]:
        m = re.findall(r"/v(\d+)", ln)
        if len(m) == 2:
            a, b = sorted(int(v) for v in m)
            edges.append([a, b])
    head = lines[0].strip()
    json.dump({"header": head, "n_nodes": int(nv.value), "adjacencies": sorted(edges), "alpha": 0.05,
               "source": "dataset/alarm/alarm_s5000_struct"},
              open(os.path.join(HERE, "alarm_s5000_struct.json"), "w"), indent=1)

    ch = {}
    for a in range(0, 9):
        for b in range(1, a + 1):
            out = np.zeros((200, b), dtype=np.int32)
            k = L.ref_choices(a, b, out.ctypes.data, 200)
            ch[f"{a},{b}"] = out[:k].tolist()
    for a, b in ((12, 3), (10, 5), (15, 2)):
        cap = 20000
        out = np.zeros((cap, b), dtype=np.int32)
        k = L.ref_choices(a, b, out.ctypes.data, cap)
        ch[f"{a},{b}"] = out[:k].tolist()
    json.dump(ch, open(os.path.join(HERE, "ref_choices.json"), "w"))

    # tables on ALARM (reference coding) and on a small synthetic sample with an odd row count
    rng = np.random.default_rng(7)
    syn = datagen.random_sample(11, 9, 1237, 2, 5)
    sets = {"alarm": (cols, nr.value, card), "syn": (syn.cols, syn.n_rows, syn.card)}
    rec = {}
    for name, (cc, nrow, cd) in sets.items():
        n = cc.shape[0]
        tests, tabs, nis, njs, nks = [], [], [], [], []
        for d in range(0, 4):
            for _ in range(12):
                vs = rng.choice(n, size=d + 2, replace=False)
                x, y = int(min(vs[0], vs[1])), int(max(vs[0], vs[1]))
                z = np.sort(vs[2:]).astype(np.int32)
                dimz = int(np.prod(cd[z])) if d else 1
                nt = np.zeros(dimz * cd[x] * cd[y], np.int32)
                ni = np.zeros(dimz * cd[x], np.int32)
                nj = np.zeros(dimz * cd[y], np.int32)
                nk = np.zeros(dimz, np.int32)
                cc_c = np.ascontiguousarray(cc)
                L.ref_counts(cc_c.ctypes.data, cc_c.shape[1], n, nrow, cd.ctypes.data, x, y,
                             z.ctypes.data if d else None, d, nt.ctypes.data, ni.ctypes.data, nj.ctypes.data,
                             nk.ctypes.data)
                tests.append([x, y, d] + z.tolist() + [-1] * (3 - d))
                tabs.append(nt)
                nis.append(ni)
                njs.append(nj)
                nks.append(nk)
        off = np.cumsum([0] + [len(t) for t in tabs])
        rec[name] = dict(tests=np.array(tests, np.int32), n=np.concatenate(tabs), n_off=off,
                         ni=np.concatenate(nis), nj=np.concatenate(njs), nk=np.concatenate(nks))
    np.savez_compressed(os.path.join(HERE, "ref_counts.npz"),
                        syn_cols=syn.cols[:, : syn.n_rows], syn_card=syn.card,
                        **{f"{k}_{f}": v for k, r in rec.items() for f, v in r.items()})

    groups, gtabs = [], []
    for _ in range(6):
        g = int(rng.integers(2, 6))
        d = int(rng.integers(1, 3))
        x, y = sorted(int(v) for v in rng.choice(nv.value, 2, replace=False))
        others = [v for v in range(nv.value) if v not in (x, y)]
        zs = []
        for _m in range(g):
            zs.extend(sorted(int(v) for v in rng.choice(others, d, replace=False)))
        z = np.array(zs, np.int32)
        total = sum(int(np.prod(card[z[m * d:(m + 1) * d]])) for m in range(g))
        nt = np.zeros(total * card[x] * card[y], np.int32)
        cc_c = np.ascontiguousarray(cols)
        L.ref_counts_group(cc_c.ctypes.data, cc_c.shape[1], nv.value, nr.value, card.ctypes.data, x, y,
                           z.ctypes.data, d, g, nt.ctypes.data)
        groups.append([x, y, d, g] + zs)
        gtabs.append(nt)
    np.savez_compressed(os.path.join(HERE, "ref_counts_group.npz"),
                        groups=np.array(json.dumps(groups)),
                        n=np.concatenate(gtabs), n_off=np.cumsum([0] + [len(t) for t in gtabs]))
    shd_fixtures(L)
    mle_fixtures(L, cols[:, : nr.value], nr.value, card)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
