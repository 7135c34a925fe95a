"""The C++ host mirror end to end on the device: ALARM through PCStable::StructLearnCompData
(CSV loader, step 1 on the GPU through the C ABI, orientation, structure file) must write
the reference's own saved CPDAG byte for byte and count the oracle's CI tests."""
import json
import os

import pytest

from oracle import oracle as O
from tests.helpers import GOLDEN, alarm_sample
from tests.test_host import alarm_csv, parse_struct, run_cli

pytestmark = pytest.mark.gpu


def test_pcstable_host_class_alarm_end_to_end(tmp_path):
    csv = alarm_csv(tmp_path)
    out = tmp_path / "alarm_s5000_struct"
    print("gpu host: ALARM N=5000 alpha 0.05 through fpgm_pcstable", flush=True)
    r = run_cli("-alpha", "0.05", "-o", out, "--stats", csv)
    st = json.loads(r.stdout.strip().splitlines()[-1])
    assert out.read_text() == open(os.path.join(GOLDEN, "alarm_s5000_struct.txt")).read()
    o = O.pc_skeleton(O.Data.of(alarm_sample()), 0.05, 1000, n_threads=8)
    assert st["num_ci_test"] == o["num_ci_test"]
    assert st["edges"] == 44


def test_pcstable_host_class_dag_and_root(tmp_path):
    csv = alarm_csv(tmp_path)
    out = tmp_path / "dag"
    print("gpu host: ALARM DAG completion with ROOT", flush=True)
    run_cli("-alpha", "0.05", "--dag", "--add-root", "-o", out, csv)
    head, und, arcs = parse_struct(out.read_text())
    assert not und
    pairs = {tuple(sorted(a)) for a in arcs if 37 not in a}
    ref = {tuple(sorted(int(v.split("/v")[1]) for v in ln.split(": ", 1)[1].replace(" -> ", " -- ").split(" -- ")))
           for ln in open(os.path.join(GOLDEN, "alarm_s5000_struct.txt")).read().strip().split("\n")[1:]}
    assert pairs == ref
    assert head.startswith("number of nodes = 38")
