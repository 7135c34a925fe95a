"""PC-stable CI-stage benchmark on one MI355X.

One step = the whole CI stage (level 0 through the last level, PCStable.cpp:109-232)
of PC-stable on a seeded synthetic workload whose uint8 columns are already resident
in HBM.  The metric is reference-equivalent CI tests per second (num_ci_test with group
size 1, speculative tests excluded) over the timed steps.

    python bench.py [--steps K] [--warmup W] [--workload 3]

Prints one JSON line.  Beside the metric it reports the plain one-launch-per-test form
(first tests of every level, sampled, the rest batched so the stage stays exact), the
counting kernel's achieved bandwidth against HBM peak, and the CPU restatement timed on
a bounded sample of the same levels on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fastpgm_amd import _lib, configs  # noqa: E402

HBM_PEAK_GBS = 8000.0


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


def cpu_baseline(sample, w, g, seconds, threads):
    """The oracle (CPU restatement) on a bounded sample of every level: the first edges of
    the level's frozen adjacency, sized to about seconds / levels each; the stage's CPU
    time is then extrapolated level by level from the measured rates."""
    from oracle import oracle as O
    D = O.Data.of(sample)
    L = len(g["tests_per_level"])
    adjs = adjacency_before_levels(sample.n_vars, g["sepsets"], L)
    per = max(0.5, seconds / max(1, sum(1 for t in g["tests_per_level"] if t > 0)))
    est, sampled, sampled_s = 0.0, 0, 0.0
    # the level with the most tests gets what the others leave of the budget
    main = int(np.argmax(g["tests_per_level"]))
    order = [l for l in range(L) if l != main] + [main]
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
        log(f"cpu level {lvl}: {k} tests on {m} edges in {dt:.2f} s ({rate:.0f} tests/s)")
    value = g["num_ci_test"] / est if est > 0 else None
    return {"value": value, "unit": "CI tests/s", "cores": threads, "kind": "port",
            "sample": f"oracle CheckEdge on the first edges of each level's frozen adjacency: {sampled} tests "
                      f"in {sampled_s:.1f} s on {threads} threads, stage time extrapolated per level "
                      f"(est. {est:.1f} s for {g['num_ci_test']} tests)"}


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
    total, kernel_ms, alg_bytes, launches = 0, 0.0, 0.0, 0
    t_start = time.perf_counter()
    for i in range(args.steps):
        g = ctx.pc_skeleton_multi(w.alpha, depth, want_adj=False)
        ks = ctx.kernel_stats()
        total += g["num_ci_test"]
        kernel_ms += ks["kernel_ms"]
        alg_bytes += ks["alg_bytes"]
        launches += ks["launches"]
    elapsed = time.perf_counter() - t_start
    value = total / elapsed
    freq = g["edge_freq"]
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {value:.0f} tests/s; per level {g['tests_per_level']}")
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_count_stat",
                     "kernel_ms_per_step": kernel_ms / args.steps, "launches_per_step": launches / args.steps,
                     "alg_bytes_per_step": alg_bytes / args.steps},
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


def hbm_traffic_from_counters(fetch_csv, write_csv, stages, copy_bytes=1 << 30):
    """HBM bytes moved by the counting kernels per CI stage, from rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE counter_collection.csv files (KiB per dispatch) of a
    bench run made with --copy-calibration that ran `stages` CI stages: each counter
    is scaled by copy_bytes over what it read for the largest copy dispatch (the
    calibration copy), then divided by the stage count."""
    import csv

    def totals(path):
        rows = list(csv.DictReader(open(path)))
        copy = max(float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"])
        kern = [float(r["Counter_Value"]) for r in rows if "k_count_stat" in r["Kernel_Name"]]
        scale = copy_bytes / (copy * 1024.0)
        return sum(kern) * 1024.0 * scale / stages, len(kern) / stages, scale

    f, n, ff = totals(fetch_csv)
    wr, _, wf = totals(write_csv)
    return {"fetch_bytes": f, "write_bytes": wr, "dispatches": n, "fetch_scale": ff, "write_scale": wf,
            "bytes_per_launch": (f + wr) / max(n, 1)}


def summarize_pmc(pmc_dir):
    """Write the counting kernels' per-stage HBM traffic into profiles/hbm_traffic.json
    from a directory holding fetch/ and write/ (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE
    output) and run.json (the JSON line the profiled bench run printed).  No GPU use."""
    import glob
    run = json.load(open(os.path.join(pmc_dir, "run.json")))
    fc = glob.glob(os.path.join(pmc_dir, "fetch", "*counter_collection.csv"))
    wc = glob.glob(os.path.join(pmc_dir, "write", "*counter_collection.csv"))
    if not fc or not wc:
        raise SystemExit(f"no counter_collection.csv under {pmc_dir}/fetch and {pmc_dir}/write")
    t = hbm_traffic_from_counters(fc[0], wc[0], run["stages_in_run"])
    rl = run["roofline"]
    alg, kms = rl["alg_bytes_per_step"], rl["kernel_ms_per_step"]
    t.update({"alg_bytes": alg, "traffic_over_alg": (t["fetch_bytes"] + t["write_bytes"]) / alg,
              "kernel_ms_step": kms,
              "hbm_gbs_measured": (t["fetch_bytes"] + t["write_bytes"]) / (kms * 1e-3) / 1e9})
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    rec = json.load(open(path)) if os.path.exists(path) else {}
    name = run["config"]["workload"]
    rec[name] = t
    with open(path, "w") as fh:
        json.dump(rec, fh, indent=1)
        fh.write("\n")
    lines = [
        f"k_count_stat -- HBM traffic, workload {name}, per CI stage",
        "Commands (GPU box, repository root, TMPDIR=/tmp), one counter per run, tuning records from a",
        "bench run with --save-tune:",
        "  rocprofv3 --pmc FETCH_SIZE -f csv -d gpurun_out/pmc_w3/fetch -o w3 -- python bench.py --steps 1 \\",
        "      --warmup 0 --no-cpu --no-plain --copy-calibration --tune-file gpurun_out/tune_w3.json",
        "  rocprofv3 --pmc WRITE_SIZE -f csv -d gpurun_out/pmc_w3/write -o w3 -- (the same bench command)",
        "  (the FETCH run's JSON line saved as gpurun_out/pmc_w3/run.json), then, on any host:",
        "  python bench.py --summarize-pmc gpurun_out/pmc_w3     -> profiles/hbm_traffic.json and this file",
        f"Each profiled run executes {run['stages_in_run']} CI stages; totals are divided by that.",
        "Calibration: --copy-calibration runs one 1 GiB device-to-device hipMemcpy (__amd_rocclr_copyBuffer);",
        f"  each counter is scaled by 1 GiB over what it read for that copy: fetch scale "
        f"{t['fetch_scale']:.4f}, write scale {t['write_scale']:.4f}.",
        f"Counting kernels (k_count_stat<0..4>), {t['dispatches']:.0f} dispatches per stage:",
        f"  fetch  {t['fetch_bytes']:.4g} bytes, write {t['write_bytes']:.4g} bytes (scaled)",
        f"  total  {t['fetch_bytes'] + t['write_bytes']:.4g} bytes = {t['traffic_over_alg']:.3f} x the algorithmic "
        f"bytes {alg:.4g}",
        f"  per dispatch {t['bytes_per_launch']:.4g} bytes",
        f"  kernel time {kms:.1f} ms per stage -> measured HBM rate {t['hbm_gbs_measured']:.0f} GB/s of 8000.",
        "Reading: the traffic is below the algorithmic bytes because the kernels read the packed",
        f"columns ({run['roofline'].get('packed', {}).get('bytes_per_step', 0):.4g} bytes per stage at 2, 4 "
        "or 8 bits per value by cardinality) and tests",
        "launched together share columns (every level-1 test of an edge scans x and y; concurrent tests",
        "share z columns): the second read of a column comes from L2 / MALL.  HBM does not bound these",
        "kernels: the per-row work (unpacking, key arithmetic and one LDS read-modify-write per row and",
        "test) does -- the SQ counter profile has the instruction mix.",
    ]
    with open(os.path.join(ROOT, "profiles", f"k_count_stat_{name}_hbm_traffic.txt"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(json.dumps({name: t}, indent=1))


def _kernel_short(name):
    return name.replace("void ", "").replace("fpgm::(anonymous namespace)::", "").split("(")[0]


def summarize_trace(trace_dir, out_path):
    """Kernel-trace profile text from DIR/*kernel_stats.csv (rocprofv3 --kernel-trace --stats) and
    DIR/run.json (the profiled run's JSON line).  No GPU use."""
    import csv
    import glob
    run = json.load(open(os.path.join(trace_dir, "run.json")))
    stats = glob.glob(os.path.join(trace_dir, "*kernel_stats.csv"))[0]
    rows = list(csv.DictReader(open(stats)))
    stages = run["stages_in_run"]
    tot = {_kernel_short(r["Name"]): (float(r["TotalDurationNs"]) / 1e6, int(r["Calls"])) for r in rows}
    counting = [k for k in tot if k.startswith(("k_count_stat", "k_fin_wide", "k_compact"))]
    cnt_ms = sum(tot[k][0] for k in counting) / stages
    cnt_calls = sum(tot[k][1] for k in counting) / stages
    rl = run["roofline"]
    gram = next((v for k, v in tot.items() if k.startswith("k_gram_l0")), (0.0, 0))
    lines = [
        f"k_count_stat / k_gram_l0 -- kernel trace, workload {run['config']['workload']}",
        "Command (GPU box, repository root, TMPDIR=/tmp; tuning records from a bench run with --save-tune):",
        "  rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_kt -o w3 -- \\",
        "      python bench.py --steps 1 --warmup 0 --no-cpu --no-plain --copy-calibration \\",
        "      --tune-file gpurun_out/tune_w3.json",
        f"  then: python bench.py --summarize-trace gpurun_out/prof_kt   (run.json = that run's JSON line)",
        f"The run executes the CI stage {stages} times (timed step + read-back run); per-stage figures are",
        "the trace totals divided by that.",
        "  (every counting kernel reads the packed columns: 2, 4 or 8 bits per value by cardinality)",
        "  k_count_stat<2>  per-lane 16-bit LDS sink, tables <= 64 cells",
        "  k_count_stat<3>  per-lane 16-bit LDS sink, tables <= 96 cells (48 KiB LDS)",
        "  k_count_stat<0>  shared LDS histogram, tables <= 8192 cells ([cell][32] columns up to 256)",
        "  k_count_stat<4>  z-sliced LDS histogram (16384-cell slices)",
        "  k_count_stat<1>  device-memory sink (larger tables)",
        "  k_fin_wide_*     grid-wide finalize of device-memory tables",
        "  k_compact        per-round removal of tests whose edge an earlier round decided",
        "  k_gram_l0        level 0 as the int8 one-hot Gram matrix (v_mfma_i32_32x32x32_i8)",
        "  k_pairs_l0, k_pack_l0  level-0 G^2 / decision and result packing",
        f"Counting launches (k_count_stat + k_fin_wide + k_compact): {cnt_ms:.2f} ms per stage summed over "
        f"{cnt_calls:.0f} dispatches per stage;",
        f"bench.py's HIP-event figure for the same run: {rl['kernel_ms_per_step']:.2f} ms per stage (elapsed",
        "  from each round's first to its last counting launch: the slice and device-memory launches,",
        "  and LDS groups smaller than the CU count, run on streams of their own beside the others, so",
        "  the trace's sum of durations exceeds the elapsed time).",
        f"Algorithmic bytes per stage (sum over executed tests of (|Z|+2) * N): {rl['alg_bytes_per_step']:.4g}",
        f"  (in the packed layout the kernels read: {rl.get('packed', {}).get('bytes_per_step', 0):.4g} bytes, "
        f"{rl.get('packed', {}).get('bytes_per_step', 0) / (cnt_ms * 1e-3) / 1e9:.0f} GB/s by the trace)",
        f"  -> achieved {rl['alg_bytes_per_step'] / (cnt_ms * 1e-3) / 1e9:.0f} GB/s by the trace, "
        f"{rl['achieved']:.0f} GB/s by the events (peak {rl['peak']:.0f}); the effective rate exceeds the",
        "     HBM peak because tests launched together share columns through L2 / MALL (measured HBM",
        "     traffic: the hbm_traffic profile).",
    ]
    if gram[1]:
        g_ms = gram[0] / gram[1]
        lines.append(f"Level 0 (k_gram_l0): {g_ms:.2f} ms per launch; 1.73e13 int8 ops per launch at "
                     f"1000 vars (W = 4075 one-hot columns padded to 4096, 528 tiles of 128 x 128, 2 N each)")
        lines.append(f"  -> {1.7301504e13 / (g_ms * 1e-3) / 1e15:.2f} POP/s of the 5.0 POP/s dense int8 peak.")
    lines += ["", open(stats).read()]
    with open(out_path, "w") as f:
        f.write("\n".join(lines))
    print("\n".join(lines[:30]))


def summarize_sq(sq_dir, out_path):
    """SQ instruction-counter profile text from DIR/*counter_collection.csv.  No GPU use."""
    import collections
    import csv
    import glob
    path = glob.glob(os.path.join(sq_dir, "*counter_collection.csv"))[0]
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
             "      SQ_INSTS_VMEM_RD SQ_LDS_BANK_CONFLICT SQ_INSTS_SALU -f csv -d gpurun_out/pmc_sq -o sq -- \\",
             "      python bench.py --steps 1 --warmup 0 --no-cpu --no-plain --tune-file gpurun_out/tune_w3.json",
             "  then: python bench.py --summarize-sq gpurun_out/pmc_sq",
             "kernel                 disp " + " ".join(f"{n[3:]:>17s}" for n in names)]
    for k in sorted(agg, key=lambda k: -agg[k]["SQ_BUSY_CYCLES"]):
        d = agg[k]
        lines.append(f"{k:22s} {len(disp[k]):4d} " + " ".join(f"{d[n]:17.4g}" for n in names))
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
    ap.add_argument("--summarize-pmc", default="", metavar="DIR",
                    help="only write profiles/hbm_traffic.json from DIR/{fetch,write}/ rocprofv3 "
                         "--pmc CSVs and DIR/run.json (the profiled run's output line); no GPU use")
    ap.add_argument("--summarize-trace", default="", metavar="DIR",
                    help="only write the kernel-trace profile from DIR (rocprofv3 --kernel-trace --stats output "
                         "and run.json); no GPU use")
    ap.add_argument("--summarize-sq", default="", metavar="DIR",
                    help="only write the SQ counter profile from DIR (rocprofv3 --pmc SQ_* output); no GPU use")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (HBM counter calibration)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes and round schedule (no fpgm_autotune before warm-up)")
    ap.add_argument("--tune-file", default="", metavar="JSON",
                    help="apply per-level tuning records saved by --save-tune instead of tuning (profiled runs)")
    ap.add_argument("--save-tune", default="", metavar="JSON", help="write the autotuner's per-level records")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace,
                        os.path.join(ROOT, "profiles", "k_count_stat_dag-1000v-1500e-1m_kernel_trace.txt"))
        return
    if args.summarize_sq:
        summarize_sq(args.summarize_sq, os.path.join(ROOT, "profiles", "k_count_stat_dag-1000v-1500e-1m_sq_counters.txt"))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the CI stage of one PC-stable run is not sharded across GPUs in "
              "this build (see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
    w = configs.get(args.workload)
    if w.n_replicates:
        bench_ensemble(args, w)
        return
    t0 = time.perf_counter()
    sample = w.sample()
    log(f"workload {w.name}: {sample.n_vars} vars x {sample.n_rows} rows generated in "
        f"{time.perf_counter() - t0:.1f} s")
    ctx = _lib.Context.of(sample)   # uploads the columns: not timed

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

    for i in range(args.warmup):
        g = ctx.pc_skeleton(w.alpha, w.depth)
        log(f"warmup {i}: {g['num_ci_test']} tests in {g['ci_seconds']:.3f} s")

    # timed region: the CI stage only (fpgm_pc_skeleton), nothing read back but the
    # adjacency and the stats struct
    total_tests, kernel_ms, alg_bytes, packed_bytes, launches = 0, 0.0, 0.0, 0.0, 0
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
    elapsed = time.perf_counter() - t_start
    value = total_tests / elapsed
    g = ctx.pc_skeleton(w.alpha, w.depth)   # untimed: full read-back for the checks below
    ref = g
    if not all(np.array_equal(a, g["adj"]) for a in adjs) or any(k != len(g["sepsets"]) for k in n_sep):
        raise RuntimeError("skeleton differs between steps")
    log(f"timed {args.steps} steps: {elapsed:.3f} s, {value:.0f} tests/s; per level "
        f"{g['tests_per_level']} executed {g['executed_per_level']}")

    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
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
        "stages_in_run": args.warmup + args.steps + 1 + extra_stages,
        "num_ci_test": g["num_ci_test"],
        "tests_per_level": g["tests_per_level"],
        "executed_per_level": g["executed_per_level"],
        "in_band_per_level": g["in_band_per_level"],
        "edges_after_level": g["edges_after_level"],
        "max_rel_err_g2_fp32_vs_fp64": g["max_rel_err_g2_fp32"],
        "untuned_stage_ms": untuned * 1e3 if untuned else None,
        "autotune": ({"seconds": tune.get("seconds"), "source": args.tune_file or "fpgm_autotune",
                      "levels": tune["levels"]} if tune else None),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_count_stat", "kernel_ms_per_step": kernel_ms / args.steps,
                     "launches_per_step": launches / args.steps,
                     "alg_bytes_per_step": alg_bytes / args.steps,
                     # the same tests in the layout the kernels read (columns packed at 2, 4 or 8
                     # bits per value by cardinality): the bytes they cannot avoid reading
                     "packed": {"bytes_per_step": packed_bytes / args.steps,
                                "achieved": packed_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0,
                                "frac": (packed_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                                         if kernel_ms > 0 else 0.0)}},
    }

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(w.name)
    if traffic:
        out["roofline"]["traffic"] = traffic["bytes_per_launch"]
        out["roofline"]["traffic_detail"] = traffic

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
