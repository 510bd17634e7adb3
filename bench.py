#!/usr/bin/env python3
"""Rule-evaluated pairs/s of the plain blocking job on one MI355X.

Workload c2: 100,000 x 100,000 synthetic rows (5 columns), one rule
year equal AND JW(title) >= 0.89 AND JW(author) >= 0.85, one partition.
A step = one blocking job over the resident tables: submit, run every
kernel, hand the matched pairs back to the caller (copied into a numpy
array).  Steps are issued with up to --inflight jobs outstanding, as the
partition pipeline issues partitions; each job still does all of its work.
Inputs are uploaded before the timed region; per-table device structures
(signatures, join index) are built by the warmup jobs, and the cold first
job and the latency of a job alone are reported separately.

Prints one JSON line.
"""
import argparse
import collections
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic int32 lane-ops per candidate pair of the bound filter (the
# dominant kernel): the test "SAD of the 32 byte-class counters > allowance"
# needs 8 v_sad_u8 (32 counters, 4 per dword) and 1 compare.  The allowance
# depends on the pair's first-bytes class and lengths; it is computed per
# run of equal B length, and A rows are ordered by first byte so that the
# class is known per 64-row slice for most B rows, so no per-pair work is
# counted for it (the per-pair class test, 4 more ops, runs only where a B
# row's first byte falls in a slice's range).
FILTER_OPS_PER_CANDIDATE = 9
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 int32 lanes/clk x 2.4 GHz
HBM_PEAK_GBS = 8000.0
# Filter-stage HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE,
# scaled by a copy kernel of known size (summary and commands in this file).
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_filter_tiles_c2_hbm_traffic.txt")


def measured_traffic(workload):
    """traffic_bytes_per_launch recorded by the PMC runs, or None."""
    if workload != "c2" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("traffic_bytes_per_launch"):
            return int(line.split("=")[1])
    return None


def measured_valu_per_candidate(workload):
    """VALU lane-instructions per candidate of k_filter_tiles recorded by the
    SQ counter run, or None."""
    path = os.path.join(ROOT, "profiles", "k_filter_tiles_c2_sq_counters.txt")
    if workload != "c2" or not os.path.exists(path):
        return None
    for line in open(path):
        if "VALU lane-instructions per candidate" in line:
            return float(line.rsplit("=", 1)[1])
    return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cfg, L, R, plan, budget_rows):
    """The CPU oracle (faithful restatement, every pair interpreted) on every
    k-th left row against all right rows: on n_threads worker threads (the
    box's CPU share: 16, or fewer cores), and on one thread over an eighth
    of that sample."""
    from oracle import oracle as O
    from tests import helpers as H
    n_threads = min(16, os.cpu_count() or 1)
    op = H.oracle_plan(plan)
    tl, tr = O.Table(L), O.Table(R)

    def run(step, threads):
        t = time.perf_counter()
        _, cnt = O.block(op, tl, tr, l_step=step, n_threads=threads)
        return cnt["pairs"], time.perf_counter() - t

    step = max(1, cfg["n_l"] // budget_rows)
    pairs, dt = run(step, n_threads)
    pairs1, dt1 = run(step * 8, 1)
    return {"value": pairs / dt, "unit": "pairs/s", "cores": n_threads,
            "kind": "port", "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "value_1_thread": pairs1 / dt1,
            "sample": "every %dth left row (%d rows) x all %d right rows = %d pairs, %.1f s on "
                      "%d threads; 1 thread: every %dth row, %d pairs, %.1f s"
                      % (step, (cfg["n_l"] + step - 1) // step, cfg["n_r"], pairs, dt,
                         n_threads, step * 8, pairs1, dt1)}


def end_to_end_cli(L, R, rules):
    """The drop-in CLI (run_hyperblocker_exec) on the same tables written as
    CSV files with the rule as YAML, one partition: the whole run in a fresh
    process (CSV load, plan, partitioning, context set-up, upload, blocking,
    output writing), with the CLI's own phase times."""
    import subprocess
    import tempfile
    from hyperblocker_amd import datagen
    exe = os.path.join(ROOT, "hyperblocker_amd", "bin", "run_hyperblocker_exec")
    if not os.path.exists(exe):
        return {"skipped": "run_hyperblocker_exec is not built"}
    with tempfile.TemporaryDirectory() as d:
        datagen.write_dataset(d, L, R, rules)
        t = time.perf_counter()
        r = subprocess.run([exe, "-data_l", d + "/tableA.csv", "-data_r", d + "/tableB.csv",
                            "-rule_dir", d + "/rules", "-o", d + "/out.csv", "-n_partitions",
                            "1", "-metrics", d + "/m.json"], stdout=subprocess.DEVNULL,
                           stderr=subprocess.PIPE, text=True, timeout=300)
        process_s = time.perf_counter() - t
        if r.returncode != 0:  # reported, not raised: the line above is still measured
            return {"error": "exit %d: %s" % (r.returncode, r.stderr.strip()[-300:])}
        m = json.load(open(d + "/m.json"))
    return {"pairs_per_s": m["pairs"] / m["total_s"], "total_s": m["total_s"],
            "load_s": m["load_s"], "partition_s": m["partition_s"],
            "blocking_s": m["blocking_s"], "process_s": process_s, "matches": m["matches"]}


def copy_calibration(n_bytes):
    """Ten device-to-device copies of n_bytes: under rocprofv3 --pmc, the
    FETCH_SIZE / WRITE_SIZE this copy kernel reports per dispatch against its
    known n_bytes gives the scale applied to the blocking kernels' counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(n_bytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(n_bytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(n_bytes)) == 0
    for _ in range(10):
        assert hip.hipMemcpy(dst, src, C.c_size_t(n_bytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": n_bytes, "copies": 10}))


def job_set_bench(args):
    """Streamed partitions (c3), batched (c4) or served (c5) job sets: one
    JSON line; c4 and c5 also time plain per-job launches on the same jobs."""
    import hyperblocker_amd as hb
    from hyperblocker_amd import workloads

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    with hb.Context(0) as ctx:
        if args.workload == "c3":
            r = workloads.run_c3(ctx, log=log)
            value, desc = r["streamed_pairs_per_s"], (
                "c3: 1M x 1M rows, 8 columns, 3 rules, %d hash partitions streamed "
                "(uploads overlapped with compute)" % r["partitions"])
            ms = r["streamed_s"] * 1e3
        elif args.workload == "c4":
            r = workloads.run_c4(ctx, n_jobs=args.jobs or 4096, log=log)
            value, desc = r["batched_pairs_per_s"], (
                "c4: %d jobs, rows per side log-uniform [100, 50k], 1-4 rules each, "
                "one batch vs one launch sequence per job" % r["jobs"])
            ms = r["batched_s"] * 1e3
        else:
            r = workloads.run_c5(ctx, n_jobs=args.jobs or 1500, log=log)
            value, desc = r["served_pairs_per_s"], (
                "c5: 2M-row resident right table, %d left jobs of 1-10k rows, 1-3 rules, "
                "Poisson arrivals at 70%% of plain capacity" % r["jobs"])
            ms = None
    out = {"metric": "rule-evaluated pairs/sec on 1 MI355X", "value": value,
           "unit": "pairs/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
           "data": "synthetic (seeded generator)", "config": {"workload": desc}}
    out.update(r)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--inflight", type=int, default=4,
                    help="jobs outstanding at once in the timed loop (a context has 4 "
                         "batch streams)")
    ap.add_argument("--cpu-rows", type=int, default=20000,
                    help="left rows in the CPU baseline sample (0 = skip)")
    ap.add_argument("--no-cli", action="store_true",
                    help="skip the end-to-end run of the CLI on CSV files")
    ap.add_argument("--jobs", type=int, default=0,
                    help="jobs of the c4/c5 job-set workloads (0 = their default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run 10 device-to-device copies of BYTES (counter scale)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py: only --gpus 1 is supported; this build measures one MI355X "
              "(across GPUs the work splits by partitions or jobs, with no exchange step)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import hyperblocker_amd as hb
    from hyperblocker_amd import datagen

    if args.workload in ("c3", "c4", "c5"):
        job_set_bench(args)
        return
    L, R, dup, cfg = datagen.make(args.workload)
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    ctx = hb.Context(0)
    ctx.synchronize()
    t = time.perf_counter()
    tl = ctx.upload(hb.ColumnTable(L))
    tr = ctx.upload(hb.ColumnTable(R))
    ctx.synchronize()
    upload_s = time.perf_counter() - t
    dp = ctx.plan(plan)

    def finish(job):
        pairs, audit, stats = job.results()
        job.release()
        return pairs, stats

    def one():
        return finish(ctx.submit(dp, tl, tr))

    t = time.perf_counter()
    pairs, cold = one()
    cold_s = time.perf_counter() - t
    for _ in range(max(0, args.warmup - 1)):
        one()
    ctx.synchronize()
    stats = []
    pending = collections.deque()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pending.append(ctx.submit(dp, tl, tr))
        if len(pending) >= max(1, args.inflight):
            pairs, st = finish(pending.popleft())
            stats.append(st)
    while pending:
        pairs, st = finish(pending.popleft())
        stats.append(st)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    # Jobs alone, one at a time: their latency, and the kernels' durations
    # undisturbed.  In the timed loop jobs in flight run on different
    # streams and overlap on the GPU, which stretches each other's kernels,
    # so the roofline is taken from these runs.
    lat, alone = [], []
    for _ in range(max(5, args.steps // 2)):
        t = time.perf_counter()
        _, st_alone = one()
        lat.append((time.perf_counter() - t) * 1e3)
        alone.append(st_alone)

    n_pairs = cfg["n_l"] * cfg["n_r"]
    ms_per_step = dt * 1e3 / args.steps
    value = n_pairs * args.steps / dt
    filt_ms = float(np.mean([s["filter_ms"] for s in alone]))
    filt_ms_timed = float(np.mean([s["filter_ms"] for s in stats]))
    cand = stats[-1]["candidates"]
    achieved = cand * FILTER_OPS_PER_CANDIDATE / (filt_ms * 1e-3) / 1e12
    traffic = measured_traffic(args.workload)
    roofline = {"bound": "valu", "kernel": "k_filter_tiles+k_filter_small",
                "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tops/s",
                "frac": achieved / VALU_PEAK_TOPS, "traffic": traffic,
                "ops_per_candidate": FILTER_OPS_PER_CANDIDATE, "candidates_per_launch": cand,
                "avg_launch_ms": filt_ms,
                "launch_ms_from": "jobs run alone after the timed loop (HIP events on the job's "
                                  "stream); with %d jobs in flight the filter stage took %.4f ms"
                                  % (args.inflight, filt_ms_timed),
                "measured_valu_instr_per_candidate": measured_valu_per_candidate(args.workload),
                "hbm_gbs": None if traffic is None else traffic / (filt_ms * 1e-3) / 1e9,
                "hbm_peak_gbs": HBM_PEAK_GBS}
    cpu = None
    if args.cpu_rows > 0:
        cpu = cpu_baseline(cfg, L, R, plan, args.cpu_rows)
    e2e = None if args.no_cli else end_to_end_cli(L, R, cfg["rules"])
    out = {
        "metric": "rule-evaluated pairs/sec on 1 MI355X",
        "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded generator, workload %s)" % args.workload,
        "config": {"workload": "%s: %d x %d rows, 1 rule (year eq, JW title>=0.89, "
                               "JW author>=0.85), n_partitions=1"
                               % (args.workload, cfg["n_l"], cfg["n_r"]),
                   "pairs_per_step": n_pairs, "matches": int(len(pairs))},
        # per job, jobs alone; in the timed loop (jobs overlapping) below
        "stage_ms": {k: float(np.mean([s[k] for s in alone]))
                     for k in ("device_ms", "join_ms", "filter_ms", "exact_ms")},
        "stage_ms_timed": {k: float(np.mean([s[k] for s in stats]))
                           for k in ("device_ms", "join_ms", "filter_ms", "exact_ms")},
        "cold_first_job_ms": cold_s * 1e3,
        "upload_ms": upload_s * 1e3,
        "pcie_inclusive_pairs_per_s": n_pairs / (upload_s + ms_per_step * 1e-3),
        "inflight": args.inflight, "job_alone_ms": float(np.median(lat)),
        "job_alone_device_ms": float(np.mean([s["device_ms"] for s in alone])),
        "filtered": stats[-1]["filtered"], "jw_exact": stats[-1]["jw_exact"],
        "roofline": roofline, "cpu_baseline": cpu,
        # run_hyperblocker_exec on the same tables as CSV files (fresh process)
        "end_to_end_cli": e2e,
    }
    print(json.dumps(out))
    dp.release()
    tl.release()
    tr.release()
    ctx.close()


if __name__ == "__main__":
    main()
