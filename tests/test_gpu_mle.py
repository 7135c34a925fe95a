"""Maximum-likelihood counts on the device (fpgm_mle_counts, through the C ABI) against the
oracle, and the host mirror's parameter learning and SHD end to end: ALARM through
fpgm_pcstable --param --ref-bif must write the reference's own parameter file byte for byte
(tests/golden/ref_alarm_param.txt.gz, from FastPGM's ParameterLearning compiled from its
sources) and print the reference's SHD."""
import gzip
import os

import numpy as np
import pytest

from fastpgm_amd import _lib, configs, datagen
from oracle import oracle as O
from tests.helpers import GOLDEN, alarm_sample
from tests.test_host import alarm_csv, run_cli

pytestmark = pytest.mark.gpu


def _check(sample, parents):
    D = O.Data.of(sample)
    co, counts, _ = O.mle(D, parents, 1)
    with _lib.Context.of(sample) as ctx:
        dco, dcounts = ctx.mle_counts(parents)
    assert np.array_equal(dco, co)
    assert np.array_equal(dcounts, counts)
    return co


def test_mle_counts_alarm_learned_parents():
    z = np.load(os.path.join(GOLDEN, "ref_mle.npz"), allow_pickle=False)
    s = alarm_sample()
    off, flat = z["alarm_par_off"], z["alarm_par"]
    print("gpu mle: ALARM, parents of the saved CPDAG", flush=True)
    _check(s, [flat[off[i]:off[i + 1]].tolist() for i in range(s.n_vars)])


def test_mle_counts_all_sink_paths():
    # per-wave LDS copies, one LDS copy and device-memory tables (15360 cells)
    z = np.load(os.path.join(GOLDEN, "ref_mle.npz"), allow_pickle=False)
    syn = datagen.sample_from_rows(z["syn_cols"].T.astype(np.int32), z["syn_card"])
    off, flat = z["syn_par_off"], z["syn_par"]
    print("gpu mle: synthetic 14 vars x 20011 rows, up to 6 parents", flush=True)
    co = _check(syn, [flat[off[i]:off[i + 1]].tolist() for i in range(syn.n_vars)])
    assert np.diff(co).max() > 8192


@pytest.mark.parametrize("n_rows", [1, 4097, 300_001])
def test_mle_counts_true_dag_parents(n_rows):
    # the generating DAG of the 300-variable workload (in-degree up to 5), several chunks
    w = configs.get(2)
    net = w.network()
    s = w.sample(n_rows)
    print(f"gpu mle: {w.name} true parents, {n_rows} rows", flush=True)
    _check(s, [list(p) for p in net.parents])


def test_mle_counts_errors_and_size_query():
    s = alarm_sample()
    with _lib.Context.of(s) as ctx:
        for bad in ([[0]] + [[]] * 2, [[99]], [[]] * (s.n_vars + 1)):
            with pytest.raises(_lib.FpgmError):
                ctx.mle_counts(bad)
        # unsorted parent lists are sorted by the binding; the ABI itself rejects them
        import ctypes as C
        off = np.array([0, 2], np.int32)
        par = np.array([5, 2], np.int32)
        co = np.zeros(2, np.int64)
        st = ctx._L.fpgm_mle_counts(ctx.h, 1, off.ctypes.data, par.ctypes.data, co.ctypes.data, None, 0)
        assert st == _lib.FPGM_EINVAL
        par = np.array([2, 5], np.int32)
        st = ctx._L.fpgm_mle_counts(ctx.h, 1, off.ctypes.data, par.ctypes.data, co.ctypes.data, None, 0)
        assert st == _lib.FPGM_OK
        assert co[1] == s.card[0] * s.card[2] * s.card[5]
        buf = np.zeros(int(co[1]) - 1, np.int32)
        st = ctx._L.fpgm_mle_counts(ctx.h, 1, off.ctypes.data, par.ctypes.data, co.ctypes.data, buf.ctypes.data,
                                    C.c_int64(len(buf)))
        assert st == _lib.FPGM_EINVAL


def test_host_parameter_learning_and_shd_alarm(tmp_path):
    csv = alarm_csv(tmp_path)
    param = tmp_path / "alarm_s5000_param"
    print("gpu host: ALARM PC-stable + SHD + parameter learning through fpgm_pcstable", flush=True)
    r = run_cli("-alpha", "0.05", "--ref-bif", os.path.join(GOLDEN, "alarm.bif"), "--param", param, csv)
    assert "SHD = 5" in r.stdout
    with gzip.open(os.path.join(GOLDEN, "ref_alarm_param.txt.gz"), "rb") as f:
        assert param.read_bytes() == f.read()
