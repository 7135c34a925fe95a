"""PC-stable CI-stage benchmark on one MI355X.

One step = the whole CI stage (level 0 through the last level, PCStable.cpp:109-232)
of PC-stable on a seeded synthetic workload whose uint8 columns are already resident
in HBM.  The metric is reference-equivalent CI tests per second (num_ci_test with group
size 1, speculative tests excluded) over the timed steps.

    python bench.py [--steps K] [--warmup W] [--workload 3]

Prints one JSON line.  Beside the metric it reports:
  roofline   the dominant counting kernel against HBM: bytes from rocprofv3 FETCH_SIZE /
             WRITE_SIZE counters read in two child runs of this bench (one CI stage each,
             before this process touches the GPU), scaled by a 1 GiB copy; the algorithmic
             rate ((|Z|+2) N bytes per test, columns shared through L2) kept apart as
             "effective"; and every sink against the resource that bounds it (LDS atomics,
             device atomics, int8 MFMA), with peaks measured by probe kernels;
  plain      the one-launch-per-test form (first tests of every level, sampled);
  cpu_baseline  the CPU restatement on a bounded sample of every level on the host cores,
             with one thread, and under the reference's 128-edge schedule;
  g2_error   FP32 against FP64 G^2 over a random sample of tests of the stage's shape;
  oracle_fixture  the stage against the oracle's committed full-size result.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time
from collections import defaultdict

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fastpgm_amd import _lib, configs  # noqa: E402

HBM_PEAK_GBS = 8000.0
INT8_PEAK_OPS = 5.0e15   # dense int8 MFMA, MI355X
SINKS = ["lane", "lane-wide", "shared", "slice", "device", "bits"]
# k_count_stat's template argument (the sink kind) -> the sink the library reports it under
SINK_OF_KIND = {"2": "lane", "3": "lane-wide", "5": "lane-wide", "0": "shared", "4": "slice", "1": "device",
                "6": "bits"}
SINK_BOUND = {"lane": ("lds_atomics", "LDS atomic adds into lane-private 16-bit counters, one per row and test"),
              "lane-wide": ("lds_atomics", "LDS atomic adds into lane-private 16-bit counters, one per row and test"),
              "shared": ("lds_atomics", "LDS atomic adds into shared histogram copies, one per row and test"),
              "slice": ("lds_atomics", "LDS atomic adds into a z slice, one per row and test"),
              "device": ("dev_atomics", "device-memory atomic adds, one per row and test"),
              "bits": ("valu", "VALU: per 32 rows an AND per (x, y) pair and an AND plus a population count per cell; "
                               "rate in row-tests/s, no probe of its own")}
TUNE_DIR = os.path.join(ROOT, "profiles")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def adjacency_before_levels(n, sepsets, n_levels):
    """Adjacency at the start of each level, rebuilt from the sepsets' removal levels."""
    full = np.ones((n, n), dtype=np.uint8) - np.eye(n, dtype=np.uint8)
    out = []
    for lvl in range(n_levels):
        a = full.copy()
        for (x, y), (rl, _z) in sepsets.items():
            if rl < lvl:
                a[x, y] = a[y, x] = 0
        out.append(a)
    return out


def cpu_info():
    """Host CPU: logical CPUs of the machine, those this process may run on, the model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


def cpu_baseline(sample, w, g, seconds, threads):
    """The oracle (CPU restatement) on a bounded sample of every level: the first edges of
    the level's frozen adjacency, sized to about seconds / levels each; the stage's CPU
    time is then extrapolated level by level from the measured rates.  On the level with
    the most tests the same edges are also run with one thread and under the reference's
    own schedule (128-edge batches of one test per edge, PCStable.cpp:275-324), and the
    stage estimate is scaled by each variant's rate there."""
    from oracle import oracle as O
    D = O.Data.of(sample)
    L = len(g["tests_per_level"])
    adjs = adjacency_before_levels(sample.n_vars, g["sepsets"], L)
    per = max(0.5, seconds / max(1, sum(1 for t in g["tests_per_level"] if t > 0)))
    est, sampled, sampled_s = 0.0, 0, 0.0
    # the level with the most tests gets what the others leave of the budget
    main = int(np.argmax(g["tests_per_level"]))
    order = [l for l in range(L) if l != main] + [main]
    main_run = None
    for lvl in order:
        tests_l = g["tests_per_level"][lvl]
        if tests_l == 0:
            continue
        n_edges = int(adjs[lvl][np.triu_indices(sample.n_vars, 1)].sum())
        m = min(n_edges, 16)
        budget = per if lvl != main else max(per, seconds - sampled_s)
        t0 = time.perf_counter()
        k, _ = O.level_sample(D, w.alpha, lvl, adjs[lvl], m, threads)
        dt = time.perf_counter() - t0
        for _ in range(3):  # the first edges need not be typical: rescale until the budget is used
            if dt >= 0.5 * budget or m >= n_edges:
                break
            m = min(n_edges, max(m + 1, int(m * budget / max(dt, 1e-4))))
            t0 = time.perf_counter()
            k, _ = O.level_sample(D, w.alpha, lvl, adjs[lvl], m, threads)
            dt = time.perf_counter() - t0
        rate = k / dt if dt > 0 and k > 0 else float("inf")
        est += tests_l / rate
        sampled += k
        sampled_s += dt
        if lvl == main:
            main_run = (m, k, dt)
        log(f"cpu level {lvl}: {k} tests on {m} edges in {dt:.2f} s ({rate:.0f} tests/s)")
    value = g["num_ci_test"] / est if est > 0 else None
    out = {"value": value, "unit": "CI tests/s", "cores": threads, "kind": "port",
           "sample": f"oracle CheckEdge on the first edges of each level's frozen adjacency: {sampled} tests "
                     f"in {sampled_s:.1f} s on {threads} threads (dynamic schedule over edges), stage time "
                     f"extrapolated per level (est. {est:.1f} s for {g['num_ci_test']} tests)",
           "host": cpu_info()}
    if main_run and main > 0 and value:
        m, k, dt = main_run
        rate_all = k / dt
        variants = {}
        # one thread: about a quarter of the budget, on the main level's first edges
        m1 = max(1, min(m, int(m * 0.25 * seconds / max(dt * threads, 1e-3))))
        t0 = time.perf_counter()
        k1, _ = O.level_sample(D, w.alpha, main, adjs[main], m1, 1)
        dt1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        kr, _ = O.level_sample_ref128(D, w.alpha, main, adjs[main], m, threads)
        dtr = time.perf_counter() - t0
        for name, kk, dd, th, what in (("threads_1", k1, dt1, 1, f"{m1} edges"),
                                       ("reference_schedule", kr, dtr, threads,
                                        f"{m} edges in 128-edge batches of one test per edge")):
            r = kk / dd if dd > 0 and kk > 0 else None
            variants[name] = {"value": value * r / rate_all if r else None, "cores": th,
                              "sample": f"level {main}: {kk} tests on {what} in {dd:.2f} s ({r:.0f} tests/s against "
                                        f"{rate_all:.0f} with {threads} threads, dynamic schedule)"}
            log(f"cpu level {main}, {name}: {kk} tests in {dd:.2f} s ({r:.0f} tests/s)")
        out.update(variants)
    return out


def bench_ensemble(args, w):
    """The bootstrap ensemble (512 replicates of one base sample, 2 000 - 50 000 rows each):
    every replicate's CI stage, level-synchronously in one context (fpgm_pc_skeleton_multi).
    The depth is bounded (--depth, default 3 = levels 0..2): duplicated rows inflate G^2 in
    bootstrap replicates, so deeper levels grow to billions of tests (2.15e9 at depth 4)."""
    from oracle import oracle as O
    depth = args.depth or 3
    base = w.sample()
    reps = list(w.replicates(base))
    rows = sum(r.n_rows for r in reps)
    log(f"workload {w.name}: {len(reps)} replicates, {rows} rows in all, depth {depth}")
    ctx = _lib.Context.of_many(reps)
    for i in range(args.warmup):
        g = ctx.pc_skeleton_multi(w.alpha, depth, want_adj=False)
        log(f"warmup {i}: {g['num_ci_test']} tests in {g['ci_seconds']:.3f} s")
    total, kernel_ms, alg_bytes, packed_bytes, launches = 0, 0.0, 0.0, 0.0, 0
    sinks = {k: defaultdict(float) for k in SINKS}
    t_start = time.perf_counter()
    for i in range(args.steps):
        g = ctx.pc_skeleton_multi(w.alpha, depth, want_adj=False)
        ks = ctx.kernel_stats()
        total += g["num_ci_test"]
        kernel_ms += ks["kernel_ms"]
        alg_bytes += ks["alg_bytes"]
        packed_bytes += ks["packed_bytes"]
        launches += ks["launches"]
        for k, v in ctx.sink_stats().items():
            for f, x in v.items():
                sinks[k][f] += x
    elapsed = time.perf_counter() - t_start
    value = total / elapsed
    freq = g["edge_freq"]
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {value:.0f} tests/s; per level {g['tests_per_level']}")
    roof = build_roofline(sinks, args.steps, None, probe_peaks(), None, kernel_ms, alg_bytes, packed_bytes, launches)
    roof["traffic_source"] = "HBM counters not read for the ensemble (its columns, 1.3 GB in all, are read through L2)"
    out = {
        "metric": "CI tests/sec (reference-equivalent, PC-stable CI stage, G2)",
        "value": value, "unit": "CI tests/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded 100-variable DAG, Dirichlet(1) CPTs, 50 000-row base sample, "
                "512 bootstrap replicates)",
        "config": {"workload": w.name, "n_vars": w.n_vars, "replicates": len(reps), "rows_total": rows,
                   "rep_rows": list(w.rep_rows), "alpha": w.alpha, "depth": depth},
        "num_ci_test": g["num_ci_test"], "tests_per_level": g["tests_per_level"],
        "executed_per_level": g["executed_per_level"],
        "edge_freq_sum": int(freq.sum()), "edge_freq_max": int(freq.max()),
        "roofline": roof,
    }
    if not args.no_plain:
        gp = ctx.pc_skeleton_multi(w.alpha, depth, mode=_lib.FPGM_MODE_PLAIN, plain_sample=args.plain_sample,
                                   want_adj=False)
        est = sum(t * ps / pt for t, pt, ps in zip(gp["tests_per_level"], gp["plain_tests_per_level"],
                                                 gp["plain_seconds_per_level"]) if t and pt)
        plain_rate = gp["num_ci_test"] / est if est > 0 else None
        same = bool((gp["edge_freq"] == freq).all())
        out["plain"] = {"value": plain_rate, "unit": "CI tests/s", "est_stage_seconds": est,
                        "sample": f"first {args.plain_sample} tests of each level, one launch + host sync each",
                        "edge_freq_identical": same}
        out["batched_over_plain"] = value / plain_rate if plain_rate else None
        log(f"plain form: {plain_rate:.0f} tests/s (sampled), identical edge frequencies: {same}")
    if not args.no_cpu:
        # the oracle on whole replicates, one after another, until about --cpu-seconds
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        k, tests, t0 = 0, 0, time.perf_counter()
        while k < len(reps) and time.perf_counter() - t0 < args.cpu_seconds:
            o = O.pc_skeleton(O.Data.of(reps[k]), w.alpha, depth, n_threads=threads)
            tests += o["num_ci_test"]
            k += 1
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": tests / dt, "unit": "CI tests/s", "cores": threads, "kind": "port",
                               "sample": f"oracle PC-stable (depth {depth}) on the first {k} replicates: "
                                         f"{tests} tests in {dt:.1f} s on {threads} threads"}
    ctx.close()
    print(json.dumps(out), flush=True)


def _kernel_short(name):
    return (name.replace("void ", "").replace("fpgm::", "").replace("(anonymous namespace)::", "")
            .split("(")[0])


def kernel_family(name):
    """The sink (or other kernel) a profiled kernel name belongs to."""
    k = _kernel_short(name)
    if k.startswith("k_count_stat<"):
        return SINK_OF_KIND.get(k[len("k_count_stat<"):].split(",")[0].strip(), "other")
    if k.startswith("k_fin_wide"):
        return "device"
    if k.startswith("k_part_keys"):  # the partitioning pass of the z-slice launch's partitioned tables
        return "slice"
    if "copyBuffer" in name:
        return "copy"
    return k.split("<")[0]


def pmc_by_family(csv_path, copy_bytes=1 << 30):
    """Bytes per kernel family from a rocprofv3 --pmc counter_collection.csv (one counter,
    KiB per dispatch), scaled by copy_bytes over what the counter read for the calibration
    copy (the largest __amd_rocclr_copyBuffer dispatch)."""
    rows = list(csv.DictReader(open(csv_path)))
    copy = max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
    scale = copy_bytes / (copy * 1024.0)
    fam = defaultdict(lambda: [0.0, 0])
    for r in rows:
        f = kernel_family(r["Kernel_Name"])
        fam[f][0] += float(r["Counter_Value"]) * 1024.0 * scale
        fam[f][1] += 1
    return {k: {"bytes": v[0], "dispatches": v[1]} for k, v in fam.items()}, scale


def tune_file_for(w):
    return os.path.join(TUNE_DIR, f"tune_{w.name}.json")


def traffic_child(args, w):
    """Child of measure_traffic, run under rocprofv3 --pmc: the calibration copy, then one CI
    stage with the committed tuning records; prints the per-sink figures as JSON."""
    copy_calibration()
    sample = w.sample()
    ctx = _lib.Context.of(sample)
    if args.tune_file and os.path.exists(args.tune_file):
        with open(args.tune_file) as f:
            ctx.tune_set(json.load(f)["levels"])
    _, st = ctx.run_stage(w.alpha, w.depth)
    out = {"stages": 1, "num_ci_test": int(st.num_ci_test), "sinks": ctx.sink_stats(), "gram": ctx.gram_stats()}
    ctx.close()
    print(json.dumps(out), flush=True)


def measure_traffic(args, w):
    """HBM bytes per kernel family of one CI stage, measured with rocprofv3 --pmc FETCH_SIZE
    and then WRITE_SIZE in two child runs of this script (--traffic-child).  Runs before this
    process makes any GPU call.  Returns (result, None) or (None, reason)."""
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None, "rocprofv3 not found"
    base = tempfile.mkdtemp(prefix="fpgm_pmc_", dir="/tmp")
    tune = tune_file_for(w)
    fams, scales, child = {}, {}, None
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(base, counter)
            cmd = [exe, "--pmc", counter, "-f", "csv", "-d", d, "-o", "run", "--", sys.executable,
                   os.path.abspath(__file__), "--traffic-child", "--workload", str(w.key)]
            if os.path.exists(tune):
                cmd += ["--tune-file", tune]
            log(f"traffic: rocprofv3 --pmc {counter}, one CI stage in a child process")
            t0 = time.perf_counter()
            p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                                 env={**os.environ, "TMPDIR": "/tmp"}, start_new_session=True)
            try:
                out, err = p.communicate(timeout=args.traffic_timeout)
            except subprocess.TimeoutExpired:
                os.killpg(p.pid, signal.SIGKILL)
                p.communicate()
                return None, f"rocprofv3 --pmc {counter} exceeded {args.traffic_timeout} s"
            lines = [ln for ln in out.splitlines() if ln.startswith("{")]
            csvs = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            # rocprofv3 may end with a fault of its own after the program has finished: the
            # program's last line and the counter file decide
            if not lines or not csvs:
                tail = (err or "").strip().splitlines()[-3:]
                return None, f"rocprofv3 --pmc {counter}: no result (rc {p.returncode}): {' | '.join(tail)}"
            child = json.loads(lines[-1])
            fams[counter], scales[counter] = pmc_by_family(csvs[0])
            log(f"traffic: {counter} done in {time.perf_counter() - t0:.1f} s")
    finally:
        shutil.rmtree(base, ignore_errors=True)
    fam = {}
    for k in set(fams["FETCH_SIZE"]) | set(fams["WRITE_SIZE"]):
        f = fams["FETCH_SIZE"].get(k, {"bytes": 0.0, "dispatches": 0})
        wr = fams["WRITE_SIZE"].get(k, {"bytes": 0.0, "dispatches": 0})
        fam[k] = {"fetch_bytes": f["bytes"], "write_bytes": wr["bytes"], "bytes": f["bytes"] + wr["bytes"],
                  "dispatches": max(f["dispatches"], wr["dispatches"])}
    return {"per_stage": fam, "fetch_scale": scales["FETCH_SIZE"], "write_scale": scales["WRITE_SIZE"],
            "child_sinks": child["sinks"], "tune_file": os.path.relpath(tune, ROOT) if os.path.exists(tune) else None,
            "source": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each over one CI stage in a child run of "
                      "this bench; every counter scaled by 1 GiB over what it read for a 1 GiB device copy"}, None


def probe_peaks():
    """Measured peaks of the resources the sinks are bound by (fastpgm_amd/libfpgm_probe.so)."""
    import ctypes as C
    path = os.path.join(ROOT, "fastpgm_amd", "libfpgm_probe.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.fpgm_probe_lds_atomics.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.fpgm_probe_dev_atomics.argtypes = [C.c_int64, C.POINTER(C.c_double)]
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    if L.fpgm_probe_lds_atomics(C.byref(a), C.byref(b)) or L.fpgm_probe_dev_atomics(1 << 20, C.byref(c)):
        return None
    return {"lds_atomics": a.value, "lds_atomics_per_cu_clock": b.value, "dev_atomics": c.value,
            "probes": "k_probe_lds_atomics (lane-private [cell][lane] words, 16 adds per step), "
                      "k_probe_dev_atomics (scattered adds into a 4 MiB table)"}


def build_roofline(sinks, steps, traffic, peaks, gram, kernel_ms, alg_bytes, packed_bytes, launches):
    """The roofline object: the dominant counting kernel against HBM with measured bytes, its
    algorithmic rate apart, and each sink against the resource that bounds it."""
    lines = {}
    for name in SINKS:
        S = sinks.get(name)
        if not S or S["launches"] == 0 or S["ms"] <= 0:
            continue
        ms = S["ms"] / steps
        sec = ms * 1e-3
        L = {"kernel": f"k_count_stat ({name} sink)", "launches": S["launches"] / steps, "ms": ms,
             "tests": S["tests"] / steps, "alg_bytes": S["alg_bytes"] / steps,
             "effective_gbs": S["alg_bytes"] / steps / sec / 1e9,
             "packed_gbs": S["packed_bytes"] / steps / sec / 1e9}
        fam = (traffic or {}).get("per_stage", {}).get(name)
        if fam:
            L["hbm_bytes"] = fam["bytes"]
            L["hbm_gbs"] = fam["bytes"] / sec / 1e9
            L["hbm_frac"] = L["hbm_gbs"] / HBM_PEAK_GBS
        res, what = SINK_BOUND[name]
        rate = S["row_tests"] / steps / sec
        peak = (peaks or {}).get(res)
        L["bound"] = {"resource": what, "achieved": rate, "unit": "adds/s", "peak": peak,
                      "frac": rate / peak if peak else None}
        lines[name] = L
    if gram and gram.get("launches"):
        g_ms = gram["gram_ms"]
        gfam = (traffic or {}).get("per_stage", {}).get("k_gram_l0")
        lines["gram"] = {"kernel": "k_gram_l0 (level 0)", "launches": gram["launches"], "ms": g_ms,
                         **({"hbm_bytes": gfam["bytes"], "hbm_gbs": gfam["bytes"] / (g_ms * 1e-3) / 1e9,
                             "hbm_frac": gfam["bytes"] / (g_ms * 1e-3) / 1e9 / HBM_PEAK_GBS} if gfam else {}),
                         "bound": {"resource": "int8 MFMA (v_mfma_i32_32x32x32_i8)",
                                   "achieved": gram["int8_ops"] / (g_ms * 1e-3), "unit": "OP/s",
                                   "peak": INT8_PEAK_OPS, "frac": gram["int8_ops"] / (g_ms * 1e-3) / INT8_PEAK_OPS}}
    dom = max((k for k in lines if k != "gram"), key=lambda k: lines[k]["ms"], default=None)
    roof = {"bound": "hbm", "kernel": lines[dom]["kernel"] if dom else "k_count_stat", "unit": "GB/s",
            "peak": HBM_PEAK_GBS, "achieved": None, "frac": None, "traffic": None}
    if dom:
        D = lines[dom]
        per_launch = D["ms"] / max(D["launches"], 1e-9)
        roof["avg_launch_ms"] = per_launch
        roof["alg_bytes_per_launch"] = D["alg_bytes"] / max(D["launches"], 1e-9)
        if "hbm_bytes" in D:
            fam = traffic["per_stage"][dom]
            roof["traffic"] = fam["bytes"] / max(fam["dispatches"], 1)  # the profiled run's own dispatches
            roof["achieved"] = D["hbm_gbs"]
            roof["frac"] = D["hbm_frac"]
        roof["effective"] = {"achieved": D["effective_gbs"], "frac": D["effective_gbs"] / HBM_PEAK_GBS,
                             "note": "algorithmic (|Z|+2) N bytes per test over the launch time; tests launched "
                                     "together read their columns through L2, so this is not an HBM rate"}
        roof["limit"] = D["bound"]
    total_hbm = sum(v["bytes"] for k, v in (traffic or {}).get("per_stage", {}).items() if k in SINKS)
    roof["stage"] = {"counting_ms": kernel_ms / steps, "launches": launches / steps,
                     "alg_bytes": alg_bytes / steps, "packed_bytes": packed_bytes / steps,
                     "hbm_bytes": total_hbm or None,
                     "hbm_gbs": total_hbm / (kernel_ms / steps * 1e-3) / 1e9 if total_hbm else None}
    roof["sinks"] = lines
    roof["traffic_source"] = (traffic or {}).get("source")
    roof["peaks"] = peaks
    return roof


def g2_error_sample(ctx, sample, w, g, n=20000, seed=7):
    """FP32 G^2 against FP64 over a random sample of tests shaped like the stage's: a level
    drawn in proportion to its executed tests, an edge of that level's frozen adjacency, a
    random conditioning set of that size from adj(x) minus y.  The same tests run once as
    the stage runs them (FP32, FP64 only inside the recheck band) and once in FP64."""
    rng = np.random.default_rng(seed)
    L = len(g["tests_per_level"])
    adjs = adjacency_before_levels(sample.n_vars, g["sepsets"], L)
    ex = np.asarray(g["executed_per_level"], dtype=np.float64)
    p = ex / ex.sum()
    edges = [np.argwhere(np.triu(a, 1)) for a in adjs]
    tests = []
    while len(tests) < n:
        d = int(rng.choice(L, p=p))
        if len(edges[d]) == 0:
            continue
        x, y = (int(v) for v in edges[d][rng.integers(len(edges[d]))])
        nb = [v for v in np.flatnonzero(adjs[d][x]) if v != y]
        if len(nb) < d:
            continue
        tests.append((x, y, tuple(sorted(int(v) for v in rng.choice(nb, d, replace=False)))))
    r32 = ctx.ci_batch(tests, alpha=w.alpha, flags=0)
    r64 = ctx.ci_batch(tests, alpha=w.alpha, flags=_lib.FPGM_F_EXACT)
    m = (~r32["exact"]) & (r64["g2"] > 0)
    diff = np.abs(r32["g2"][m] - r64["g2"][m])
    rel = diff / r64["g2"][m]
    big = r64["g2"][m] >= 1.0
    crit = np.array([_lib.chisq_crit(int(df), w.alpha) for df in r64["df"][m]]) if rel.size else np.zeros(0)
    return {"tests": n, "fp32_decided": int(m.sum()), "in_band": int(r32["in_band"].sum()),
            "max_rel_err": float(rel.max()) if rel.size else 0.0, "mean_rel_err": float(rel.mean()) if rel.size else 0.0,
            "p99_rel_err": float(np.quantile(rel, 0.99)) if rel.size else 0.0,
            "max_rel_err_g2_ge_1": float(rel[big].max()) if big.any() else 0.0,
            "max_abs_err": float(diff.max()) if rel.size else 0.0,
            "max_err_over_crit": float((diff / crit).max()) if rel.size else 0.0,
            "note": "relative errors of G^2 near 0 are cancellation in the FP32 sum (absolute error about "
                    "2^-24 of the summed |terms|); the decision-relevant figure is the error over the critical "
                    "value, against the +-2 % recheck band",
            "decisions_identical": bool(np.array_equal(r32["indep"], r64["indep"])),
            "sample": "levels in proportion to executed tests, random edge and conditioning set per test"}


def oracle_fixture_check(w, sample, g):
    """The stage against the oracle's committed result for this workload (tests/golden), when
    the regenerated data has the fixture's SHA-256."""
    path = os.path.join(ROOT, "tests", "golden", f"stage_{w.name}.npz")
    if not os.path.exists(path):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import sepset_digest
    with np.load(path, allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    rel = os.path.relpath(path, ROOT)
    if sample.sha256() != str(f["data_sha256"]):
        return {"identical": False, "fixture": rel, "reason": "data SHA-256 differs from the fixture's"}
    n = sample.n_vars
    edges = np.array([(i, j) for i in range(n) for j in range(i + 1, n) if g["adj"][i, j]], dtype=np.int32)
    same = (g["tests_per_level"] == f["tests_per_level"].tolist()
            and g["edges_after_level"] == f["edges_after_level"].tolist()
            and g["num_ci_test"] == int(f["num_ci_test"])
            and np.array_equal(edges.reshape(-1, 2), f["edges"])
            and sepset_digest(g["sepsets"]) == str(f["sep_digest"]))
    return {"identical": bool(same), "fixture": rel,
            "checked": "tests and edges per level, num_ci_test, final skeleton, SHA-256 of every sepset record"}


def summarize_trace(trace_dir, out_path):
    """Kernel-trace profile text from DIR/*kernel_stats.csv (rocprofv3 --kernel-trace --stats) and
    DIR/run.json (the profiled run's JSON line, the same command without the profiler is the bench
    itself).  No GPU use."""
    import glob as _glob
    run = json.load(open(os.path.join(trace_dir, "run.json")))
    stats = _glob.glob(os.path.join(trace_dir, "**", "*kernel_stats.csv"), recursive=True)[0]
    rows = list(csv.DictReader(open(stats)))
    stages = run["stages_in_run"]
    rl = run["roofline"]
    fam = defaultdict(lambda: [0.0, 0])
    for r in rows:
        f = kernel_family(r["Name"])
        fam[f][0] += float(r["TotalDurationNs"]) / 1e6
        fam[f][1] += int(r["Calls"])
    lines = [
        f"k_count_stat / k_gram_l0 -- kernel trace, workload {run['config']['workload']}",
        "Command (GPU box, repository root, TMPDIR=/tmp):",
        f"  rocprofv3 --kernel-trace --stats -f csv -d {trace_dir} -o w3 -- python bench.py --steps 1 \\",
        "      --warmup 0 --no-cpu --no-plain --no-traffic --no-g2 --tune-file profiles/tune_dag-1000v-1500e-1m.json",
        f"  (its JSON line saved as {trace_dir}/run.json), then, on any host:",
        f"  python bench.py --summarize-trace {trace_dir}",
        f"The run executes the CI stage {stages} times (timed step + read-back run); per-stage figures",
        "below are the trace totals divided by the stage count.",
        "Sinks (the counting kernels read the packed columns, 2, 4 or 8 bits per value, or the bit planes):",
        "  lane       k_count_stat<2>  per-lane 16-bit LDS counters, tables <= 64 cells",
        "  lane-wide  k_count_stat<3>, <5>  the same with 48 / 64 KiB of counters, <= 96 / 128 cells",
        "  shared     k_count_stat<0>  shared LDS histogram, tables <= 8192 cells",
        "  slice      k_count_stat<4>  z-sliced LDS histogram (16384-cell slices); tables of more slices",
        "             than the slice limit partitioned by slice first (k_part_keys, counted here)",
        "  device     k_count_stat<1> + k_fin_wide_*  device-memory table (beyond 4096 slices)",
        "  bits       k_count_stat<6>  |Z| = 1 tables <= 48 cells from bit planes (AND + population count)",
        "  k_gram_l0  level 0 as the int8 one-hot Gram matrix (v_mfma_i32_32x32x32_i8)",
        "",
        "family        calls/stage   ms/stage   avg us/call   bench (HIP events): ms/stage  launches/stage",
    ]
    for k in SINKS + ["k_gram_l0", "k_pairs_l0", "k_compact", "k_pvalue"]:
        if k not in fam:
            continue
        ms, calls = fam[k]
        b = rl["sinks"].get(k) or (rl["sinks"].get("gram") if k == "k_gram_l0" else None)
        bench_col = f"{b['ms']:10.2f}  {b['launches']:8.1f}" if b else ""
        lines.append(f"{k:12s} {calls / stages:10.1f} {ms / stages:10.2f} {ms / max(calls, 1) * 1e3:12.1f}   {bench_col}")
    dom = next((k for k in SINKS if rl["sinks"].get(k, {}).get("kernel") == rl["kernel"]), None)
    if dom and dom in fam:
        ms, calls = fam[dom]
        lines += ["",
                  f"Dominant kernel {rl['kernel']}: average launch {ms / calls:.3f} ms by the trace, "
                  f"{rl['avg_launch_ms']:.3f} ms by the bench's HIP events;",
                  f"  algorithmic bytes per launch {rl['alg_bytes_per_launch']:.4g} -> effective "
                  f"{rl['alg_bytes_per_launch'] / (ms / calls * 1e-3) / 1e9:.0f} GB/s by the trace (columns shared "
                  "through L2: not an HBM rate);",
                  f"  bound: {rl['limit']['resource']}: {rl['limit']['achieved']:.4g} adds/s of a measured "
                  f"{rl['limit']['peak']:.4g} (frac {rl['limit']['frac']:.3f})."]
    g = rl["sinks"].get("gram")
    if g and "k_gram_l0" in fam:
        ms, calls = fam["k_gram_l0"]
        ops = g["bound"]["achieved"] * g["ms"] * 1e-3
        lines.append(f"Level 0 (k_gram_l0): {ms / calls:.2f} ms per launch, {ops:.4g} int8 operations -> "
                     f"{ops / (ms / calls * 1e-3) / 1e15:.2f} POP/s of the 5.0 POP/s dense int8 peak.")
    lines += ["", open(stats).read()]
    with open(out_path, "w") as f:
        f.write("\n".join(lines))
    print("\n".join(lines[:40]))


def summarize_traffic(bench_json, out_path):
    """HBM-traffic profile text from a bench run's JSON line (its roofline was read from
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in child runs of the same command).  No GPU use."""
    run = json.load(open(bench_json))
    rl = run["roofline"]
    lines = [
        f"k_count_stat -- HBM traffic by counters, workload {run['config']['workload']}, per CI stage",
        "Command (GPU box, repository root): python bench.py   (the default run: before touching the GPU",
        "  it runs, as child processes, for counter in FETCH_SIZE WRITE_SIZE:",
        "    rocprofv3 --pmc $counter -f csv -d /tmp/fpgm_pmc_*/$counter -o run -- python bench.py \\",
        "        --traffic-child --workload 3 --tune-file profiles/tune_dag-1000v-1500e-1m.json",
        "  each a 1 GiB device copy and one CI stage; counters in KiB per dispatch, scaled by 1 GiB over",
        "  what they read for the copy);  then: python bench.py --summarize-traffic BENCH_JSON",
        f"Source: {rl.get('traffic_source')}",
        "",
        "sink         ms/stage  launches  HBM bytes    HBM GB/s  frac of 8000   effective GB/s   bound: achieved / peak (frac)",
    ]
    for k, L in rl["sinks"].items():
        b = L["bound"]
        hb = L.get("hbm_bytes")
        bound = (f"{b['achieved']:.3g} / {b['peak']:.3g} {b['unit']} ({b['frac']:.3f})" if b.get("peak")
                 else f"{b['achieved']:.3g} {b['unit']} (VALU issue: see the SQ-counter profile)")
        lines.append(f"{k:10s} {L['ms']:9.2f} {L['launches']:8.1f}  "
                     f"{(f'{hb:10.4g}' if hb else '         -')}  {(L.get('hbm_gbs') or 0):9.0f}  "
                     f"{(L.get('hbm_frac') or 0):11.3f}  {L.get('effective_gbs', 0):14.0f}   {bound}")
    st = rl["stage"]
    lines += ["",
              f"Counting kernels per stage: {st['counting_ms']:.1f} ms elapsed (events), {st['launches']:.0f} launches,",
              f"  algorithmic bytes {st['alg_bytes']:.4g}, packed-column bytes {st['packed_bytes']:.4g},",
              f"  measured HBM bytes {st['hbm_bytes']:.4g} -> {st['hbm_gbs']:.0f} GB/s over the counting time.",
              f"Dominant kernel {rl['kernel']}: {rl['traffic']:.4g} HBM bytes per launch, {rl['avg_launch_ms']:.3f} ms per "
              f"launch -> {rl['achieved']:.0f} GB/s = {rl['frac']:.3f} of the HBM peak;",
              f"  its own bound: {rl['limit']['resource']}, {rl['limit']['frac']:.3f} of the measured peak.",
              f"Peaks measured by the probes: {json.dumps(rl['peaks'])}"]
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def summarize_sq(sq_dir, out_path, trace_dir=None):
    """SQ instruction-counter profile text from DIR/*counter_collection.csv.  No GPU use."""
    import collections
    import csv
    import glob
    path = glob.glob(os.path.join(sq_dir, "**", "*counter_collection.csv"), recursive=True)[0]
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for r in csv.DictReader(open(path)):
        k = _kernel_short(r["Kernel_Name"])
        agg[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[k].add(r["Dispatch_Id"])
    names = ["SQ_WAVES", "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS",
             "SQ_INSTS_VMEM_RD", "SQ_LDS_BANK_CONFLICT"]
    lines = ["kernel_stats of SQ counters (totals over all dispatches of each kernel)",
             "Command (GPU box, repository root, TMPDIR=/tmp):",
             "  rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_LDS SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES \\",
             f"      SQ_INSTS_VMEM_RD SQ_LDS_BANK_CONFLICT SQ_INSTS_SALU -f csv -d {sq_dir} -o sq -- \\",
             "      python bench.py --steps 1 --warmup 0 --no-cpu --no-plain --no-traffic --no-g2 \\",
             "      --tune-file profiles/tune_dag-1000v-1500e-1m.json",
             f"  (its JSON line saved as {sq_dir}/run.json), then, with the kernel-trace run of the",
             f"  same command in {trace_dir or 'DIR'}:",
             f"  python bench.py --summarize-sq {sq_dir} --sq-times {trace_dir or 'DIR'}",
             "kernel                 disp " + " ".join(f"{n[3:]:>17s}" for n in names)]
    for k in sorted(agg, key=lambda k: -agg[k]["SQ_BUSY_CYCLES"]):
        d = agg[k]
        lines.append(f"{k:22s} {len(disp[k]):4d} " + " ".join(f"{d[n]:17.4g}" for n in names))
    lines += ["", "kernel                 VALU per LDS instr   LDS bank-conflict cycles per LDS instr"]
    for k in sorted(agg, key=lambda k: -agg[k]["SQ_BUSY_CYCLES"]):
        d = agg[k]
        if d["SQ_INSTS_LDS"] > 0:
            lines.append(f"{k:22s} {d['SQ_INSTS_VALU'] / d['SQ_INSTS_LDS']:18.2f}   "
                         f"{d['SQ_LDS_BANK_CONFLICT'] / d['SQ_INSTS_LDS']:10.3f}")
    # issue rates of the counting kernels against the CU's peaks: one wave64 VALU instruction
    # per clock (4 SIMD16s, 4 clocks each) and one wave64 LDS instruction per 4 clocks (16
    # lanes per clock), 256 CUs at 2.4 GHz, over the kernels' durations in a kernel-trace
    # run of the same command (counter collection itself slows the kernels)
    run_path = os.path.join(sq_dir, "run.json")
    stats = glob.glob(os.path.join(trace_dir or "", "**", "*kernel_stats.csv"), recursive=True) if trace_dir else []
    if os.path.exists(run_path) and stats and os.path.exists(os.path.join(trace_dir, "run.json")):
        sq_stages = json.load(open(run_path)).get("stages_in_run", 1)
        kt_stages = json.load(open(os.path.join(trace_dir, "run.json"))).get("stages_in_run", 1)
        dur = collections.defaultdict(float)
        for r in csv.DictReader(open(stats[0])):
            dur[_kernel_short(r["Name"])] += float(r["TotalDurationNs"]) * 1e-9 / kt_stages
        cu_clk = 256 * 2.4e9
        lines += ["", f"Issue rates per CU and clock (counters per stage over the kernel-trace durations per stage, "
                      f"{trace_dir}):",
                  "kernel                 s/stage    VALU instr/clk (peak 1)   LDS instr x 4 clk/clk (peak 1)"]
        for k in sorted(agg, key=lambda k: -agg[k]["SQ_BUSY_CYCLES"]):
            if not k.startswith("k_count_stat") or dur.get(k, 0.0) <= 0.0:
                continue
            d = agg[k]
            sec = dur[k]
            lines.append(f"{k:22s} {sec:8.4f}   {d['SQ_INSTS_VALU'] / sq_stages / (sec * cu_clk):22.3f}   "
                         f"{4 * d['SQ_INSTS_LDS'] / sq_stages / (sec * cu_clk):28.3f}")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def copy_calibration(n_bytes=1 << 30):
    """One device-to-device copy of n_bytes (the runtime's copy kernel), so that HBM
    counters read under rocprofv3 can be scaled by what a plain copy of known size reads."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n_bytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(n_bytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n_bytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, C.c_size_t(n_bytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {n_bytes} bytes device to device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=3)
    ap.add_argument("--plain-sample", type=int, default=2000, help="plain-form tests per level")
    ap.add_argument("--depth", type=int, default=0, help="depth limit (bootstrap ensemble only; default 3)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure HBM bytes")
    ap.add_argument("--traffic-timeout", type=float, default=240.0)
    ap.add_argument("--no-g2", action="store_true", help="skip the G^2 FP32-vs-FP64 sample (profiled runs)")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--summarize-trace", default="", metavar="DIR",
                    help="only write the kernel-trace profile from DIR (rocprofv3 --kernel-trace --stats output "
                         "and run.json); no GPU use")
    ap.add_argument("--summarize-traffic", default="", metavar="JSON",
                    help="only write the HBM-traffic profile from a bench run's JSON line; no GPU use")
    ap.add_argument("--summarize-sq", default="", metavar="DIR",
                    help="only write the SQ counter profile from DIR (rocprofv3 --pmc SQ_* output); no GPU use")
    ap.add_argument("--sq-times", default="", metavar="DIR",
                    help="with --summarize-sq: kernel-trace run of the same command whose durations the issue "
                         "rates divide by")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (HBM counter calibration)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes and round schedule (no fpgm_autotune before warm-up)")
    ap.add_argument("--tune-file", default="", metavar="JSON",
                    help="apply per-level tuning records saved by --save-tune instead of tuning (profiled runs)")
    ap.add_argument("--save-tune", default="", metavar="JSON", help="write the autotuner's per-level records")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(args.summarize_trace,
                        os.path.join(ROOT, "profiles", "k_count_stat_dag-1000v-1500e-1m_kernel_trace.txt"))
        return
    if args.summarize_traffic:
        summarize_traffic(args.summarize_traffic,
                          os.path.join(ROOT, "profiles", "k_count_stat_dag-1000v-1500e-1m_hbm_traffic.txt"))
        return
    if args.summarize_sq:
        summarize_sq(args.summarize_sq, os.path.join(ROOT, "profiles", "k_count_stat_dag-1000v-1500e-1m_sq_counters.txt"),
                     args.sq_times or None)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the CI stage of one PC-stable run is not sharded across GPUs in "
              "this build (see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    w = configs.get(args.workload)
    if args.traffic_child:
        traffic_child(args, w)
        return
    # HBM bytes by counters, in child processes, before this process makes any GPU call
    traffic, traffic_err = None, "not measured (--no-traffic)"
    if not args.no_traffic and not w.n_replicates:
        traffic, traffic_err = measure_traffic(args, w)
        if traffic_err:
            log(f"traffic: {traffic_err}")

    if args.copy_calibration:
        copy_calibration()
    if w.n_replicates:
        bench_ensemble(args, w)
        return
    t0 = time.perf_counter()
    sample = w.sample()
    log(f"workload {w.name}: {sample.n_vars} vars x {sample.n_rows} rows generated in "
        f"{time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    ctx = _lib.Context.of(sample)   # uploads the columns: reported apart, not in the timed steps
    upload_s = time.perf_counter() - t0
    log(f"context: {sample.cols.nbytes / 1e9:.2f} GB of columns uploaded, packed and scanned in {upload_s:.3f} s")

    # launch shapes and round schedule: the autotuner (after two untuned reference stages),
    # saved records, or the defaults
    tune, untuned, extra_stages = None, None, 0
    if args.tune_file:
        with open(args.tune_file) as f:
            tune = json.load(f)
        ctx.tune_set(tune["levels"])
        log(f"tuning records from {args.tune_file}")
    elif not args.no_autotune:
        g = ctx.pc_skeleton(w.alpha, w.depth)
        untuned = min(g["ci_seconds"], ctx.pc_skeleton(w.alpha, w.depth)["ci_seconds"])
        extra_stages = 2
        log(f"untuned stage: {g['num_ci_test']} tests in {untuned:.3f} s")
        tune = ctx.autotune(w.alpha, w.depth)
        log(f"autotune: {tune['seconds']:.1f} s, per level default -> tuned ms: " +
            ", ".join(f"{r['level']}: {r['default_ms']:.1f} -> {r['tuned_ms']:.1f}" for r in tune["levels"]))
        if args.save_tune:
            with open(args.save_tune, "w") as f:
                json.dump({"workload": w.name, **tune}, f, indent=1)
                f.write("\n")

    for i in range(args.warmup):
        g = ctx.pc_skeleton(w.alpha, w.depth)
        log(f"warmup {i}: {g['num_ci_test']} tests in {g['ci_seconds']:.3f} s")

    # timed region: the CI stage only (fpgm_pc_skeleton), nothing read back but the
    # adjacency and the stats struct
    total_tests, kernel_ms, alg_bytes, packed_bytes, launches = 0, 0.0, 0.0, 0.0, 0
    sinks = {k: defaultdict(float) for k in SINKS}
    gram = defaultdict(float)
    adjs = [np.zeros((sample.n_vars, sample.n_vars), dtype=np.uint8) for _ in range(args.steps)]
    n_sep = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        _, st = ctx.run_stage(w.alpha, w.depth, adj=adjs[i])
        ks = ctx.kernel_stats()
        total_tests += st.num_ci_test
        kernel_ms += ks["kernel_ms"]
        alg_bytes += ks["alg_bytes"]
        packed_bytes += ks["packed_bytes"]
        launches += ks["launches"]
        n_sep.append(st.n_sepsets)
        for k, v in ctx.sink_stats().items():
            for f, x in v.items():
                sinks[k][f] += x
        for f, x in ctx.gram_stats().items():
            gram[f] += x
    elapsed = time.perf_counter() - t_start
    value = total_tests / elapsed
    g = ctx.pc_skeleton(w.alpha, w.depth)   # untimed: full read-back for the checks below
    ref = g
    if not all(np.array_equal(a, g["adj"]) for a in adjs) or any(k != len(g["sepsets"]) for k in n_sep):
        raise RuntimeError("skeleton differs between steps")
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {value:.0f} tests/s; per level "
        f"{g['tests_per_level']} executed {g['executed_per_level']}")
    fixture = oracle_fixture_check(w, sample, g)
    if fixture:
        log(f"oracle fixture {fixture['fixture']}: identical {fixture['identical']}")
    g2err = None if args.no_g2 else g2_error_sample(ctx, sample, w, g)
    if g2err:
        log(f"G^2 FP32 vs FP64 on {g2err['tests']} sampled tests: max rel {g2err['max_rel_err']:.2e}, "
            f"decisions identical {g2err['decisions_identical']}")
    peaks = probe_peaks()
    gram_step = {k: v / args.steps for k, v in gram.items()}
    gram_step["launches"] = gram["launches"] / args.steps
    roof = build_roofline(sinks, args.steps, traffic, peaks, gram_step, kernel_ms, alg_bytes, packed_bytes, launches)
    if traffic is None:
        roof["traffic_source"] = traffic_err
    if traffic and traffic.get("tune_file"):
        roof["traffic_tune_file"] = traffic["tune_file"]
    log(f"roofline: {roof['kernel']} at {roof['achieved'] or 0:.0f} GB/s measured HBM "
        f"(frac {roof['frac'] or 0:.3f}), effective {roof.get('effective', {}).get('achieved', 0):.0f} GB/s; "
        f"bound {roof.get('limit', {}).get('resource')}: frac {roof.get('limit', {}).get('frac')}")

    out = {
        "metric": "CI tests/sec (reference-equivalent, PC-stable CI stage, G2)",
        "value": value,
        "unit": "CI tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded random DAG, Dirichlet CPTs, forward-sampled uint8 columns)",
        "config": {"workload": w.name, "n_vars": w.n_vars, "n_rows": w.n_rows, "alpha": w.alpha,
                   "depth": w.depth, "dag_edges": w.n_edges, "max_in_degree": w.max_in,
                   "card_law": list(w.card_law), "dirichlet": w.conc, "data_sha256": sample.sha256()[:16]},
        "pc_stable_step1_seconds": elapsed / args.steps,
        "upload": {"seconds": upload_s, "bytes": int(sample.cols.nbytes),
                   "note": "fpgm_ctx_create from host memory: chunked copy through pinned staging, overlapped "
                           "with packing and the value scan; outside the timed steps"},
        "stages_in_run": args.warmup + args.steps + 1 + extra_stages,
        "num_ci_test": g["num_ci_test"],
        "tests_per_level": g["tests_per_level"],
        "executed_per_level": g["executed_per_level"],
        "in_band_per_level": g["in_band_per_level"],
        "edges_after_level": g["edges_after_level"],
        "max_rel_err_g2_fp32_vs_fp64_in_band": g["max_rel_err_g2_fp32"],
        "g2_error": g2err,
        "oracle_fixture": fixture,
        "untuned_stage_ms": untuned * 1e3 if untuned else None,
        "autotune": ({"seconds": tune.get("seconds"), "source": args.tune_file or "fpgm_autotune",
                      "levels": tune["levels"]} if tune else None),
        "roofline": roof,
    }

    if not args.no_plain:
        gp = ctx.pc_skeleton(w.alpha, w.depth, mode=_lib.FPGM_MODE_PLAIN, plain_sample=args.plain_sample)
        same = np.array_equal(gp["adj"], ref["adj"]) and gp["sepsets"] == ref["sepsets"]
        est = 0.0
        for lvl, tests_l in enumerate(gp["tests_per_level"]):
            pt, ps = gp["plain_tests_per_level"][lvl], gp["plain_seconds_per_level"][lvl]
            if tests_l and pt:
                est += tests_l * ps / pt
        plain_rate = gp["num_ci_test"] / est if est > 0 else None
        out["plain"] = {"value": plain_rate, "unit": "CI tests/s", "est_stage_seconds": est,
                        "sampled_tests": sum(gp["plain_tests_per_level"]),
                        "sample": f"first {args.plain_sample} tests of each level, one launch + host sync each",
                        "skeleton_identical": bool(same)}
        out["batched_over_plain"] = value / plain_rate if plain_rate else None
        log(f"plain form: {plain_rate:.0f} tests/s (sampled), identical skeleton: {same}")

    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(sample, w, ref, args.cpu_seconds, threads)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
