"""The C ABI library without a GPU: it loads, exports every declared symbol, and its
host-side chi-square helpers agree with scipy.  No device call is made here."""
import os
import re
import subprocess

import pytest
from scipy import stats

from fastpgm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "fpgm_ci.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fpgm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    assert sorted(names) == sorted(_lib.EXPORTS)


def test_product_library_does_not_contain_the_oracle():
    out = subprocess.run(["nm", "-D", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in out and "ref_" not in out
    ldd = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in ldd and "libfpgm_ref" not in ldd


@pytest.mark.parametrize("df", [1, 2, 5, 40, 700, 4096])
def test_host_pchisq_and_crit(df):
    for q in [0.01, 0.5, 0.95, 0.99]:
        x = stats.chi2.ppf(q, df)
        assert _lib.pchisq(x, df) == pytest.approx(q, rel=1e-10)
    for alpha in (0.05, 0.01):
        assert _lib.chisq_crit(df, alpha) == pytest.approx(stats.chi2.isf(alpha, df), rel=1e-10)
