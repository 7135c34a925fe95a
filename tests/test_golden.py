"""The oracle against vectors written by FastPGM's own compiled code (tests/golden/).

ref_choices.json: ChoiceGenerator::Next series (ChoiceGenerator.cpp:14-85).
ref_counts.npz: Counts2D / Counts3D tables and marginals (CellTable.cpp:23-91,174-291,430-455).
ref_counts_group.npz: Counts3DGroup tables (CellTable.cpp:93-172,293-427).
alarm_s5000.npz: the reference's CSV loader coding of dataset/alarm/alarm_s5000.
"""
import itertools
import json
import os

import numpy as np

from fastpgm_amd import datagen
from oracle import oracle as O

from helpers import GOLDEN, alarm_sample


def _load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_choice_generator_series():
    with open(os.path.join(GOLDEN, "ref_choices.json")) as f:
        ch = json.load(f)
    assert len(ch) > 30
    for key, seq in ch.items():
        a, b = (int(v) for v in key.split(","))
        exp = [list(c) for c in itertools.combinations(range(a), b)]
        assert seq == exp, key
        assert O.enumerate_choices(a, b).tolist() == seq


def _check_counts(z, prefix, sample):
    D = O.Data.of(sample)
    tests = z[f"{prefix}_tests"]
    off = z[f"{prefix}_n_off"]
    n_all = z[f"{prefix}_n"]
    ni_all, nj_all, nk_all = z[f"{prefix}_ni"], z[f"{prefix}_nj"], z[f"{prefix}_nk"]
    pi = pj = pk = 0
    for k, row in enumerate(tests):
        x, y, d = int(row[0]), int(row[1]), int(row[2])
        zz = [int(v) for v in row[3:3 + d]]
        t = O.counts(D, x, y, zz)
        assert np.array_equal(t.ravel(), n_all[off[k]:off[k + 1]])
        dimz, cx, cy = t.shape
        ni = ni_all[pi:pi + dimz * cx].reshape(dimz, cx)
        nj = nj_all[pj:pj + dimz * cy].reshape(dimz, cy)
        pi += dimz * cx
        pj += dimz * cy
        assert np.array_equal(ni, t.sum(axis=2)) and np.array_equal(nj, t.sum(axis=1))
        if d > 0:
            assert np.array_equal(nk_all[pk:pk + dimz], t.sum(axis=(1, 2)))
        pk += dimz
    assert len(tests) == 48


def test_reference_counts_alarm():
    _check_counts(_load("ref_counts.npz"), "alarm", alarm_sample())


def test_reference_counts_synthetic_odd_rows():
    z = _load("ref_counts.npz")
    s = datagen.sample_from_rows(z["syn_cols"].T.astype(np.int32), z["syn_card"])
    assert s.n_rows == 1237
    _check_counts(z, "syn", s)


def test_reference_group_counts_equal_single_counts():
    z = _load("ref_counts_group.npz")
    groups = json.loads(str(z["groups"]))
    s = alarm_sample()
    D = O.Data.of(s)
    for k, g in enumerate(groups):
        x, y, d, gs = g[:4]
        zs = g[4:]
        tabs = [O.counts(D, x, y, zs[m * d:(m + 1) * d]).ravel() for m in range(gs)]
        assert np.array_equal(np.concatenate(tabs), z["n"][z["n_off"][k]:z["n_off"][k + 1]])


def test_alarm_fixture_shape():
    s = alarm_sample()
    assert s.n_vars == 37 and s.n_rows == 5000
    assert s.card.min() >= 2 and s.card.max() <= 4
