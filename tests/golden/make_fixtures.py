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
    return L


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
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
