#!/usr/bin/env python3
"""Rule-evaluated pairs/s of the plain blocking job on one MI355X.

Workload c2: 100,000 x 100,000 synthetic rows (5 columns), one rule
year equal AND JW(title) >= 0.89 AND JW(author) >= 0.85, one partition.
The inputs are both tables in the reference's row-major layout
(SerializedTable), resident in device memory before the timed region, as
blocking_kernel receives them.  A step = one fresh blocking job: two new
tables unpacked from those rows (hb_table_from_device), every per-table
structure built (dictionary codes, signatures, join layouts), every kernel
of the job, and the matched pairs handed back to the caller (copied into a
numpy array).  Steps are issued with up to --inflight jobs outstanding, as
the partition pipeline issues partitions: one thread submits while another
collects results in order (the reference's producer and reducer threads).

Also reported: the same steps with the rows handed over from host memory
(PCIe-inclusive), the repeated-job rate of one pair of warm resident tables,
per-stage times of jobs alone, the served-vs-plain trace of the c5 workload
(p50/p99 latency), the CPU oracle on a sample, and the drop-in CLI end to
end.  Prints one JSON line.
"""
import argparse
import collections
import json
import os
import queue
import sys
import threading
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
            "blocking_s": m["blocking_s"], "ctx_wait_s": m["ctx_wait_s"],
            "upload_s": m["upload_s"], "write_s": m["write_s"], "process_s": process_s,
            "matches": m["matches"]}


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
            r = workloads.run_c3(ctx, reps=args.reps, log=log)
            value, desc = r["serialized_streamed_pairs_per_s"], (
                "c3: 1M x 1M rows, 8 columns, 3 rules, %d hash partitions streamed from "
                "row-major serialized tables (uploads overlapped with compute)" % r["partitions"])
            ms = r["serialized_streamed_s"] * 1e3
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
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served-vs-plain trace (c5) in the default run")
    ap.add_argument("--serve-jobs", type=int, default=1500,
                    help="jobs of the served trace in the default run")
    ap.add_argument("--reps", type=int, default=2,
                    help="timed passes over c3's partitions per handoff (after one untimed)")
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
    n_pairs = cfg["n_l"] * cfg["n_r"]
    ctx = hb.Context(0)
    ctx.synchronize()
    dp = ctx.plan(plan)
    # The inputs: both tables in the reference's row-major layout
    # (SerializedTable), resident in device memory before timing, as
    # blocking_kernel receives them.
    sl, sr = hb.SerializedTable(L), hb.SerializedTable(R)
    dl, dr = hb.DeviceSerializedTable(sl), hb.DeviceSerializedTable(sr)

    def finish(item):
        job, tables = item
        pairs, audit, stats = job.results()
        job.release()
        for t in tables:
            t.release()
        return pairs, stats

    def fresh():
        """One step: new tables unpacked from the device rows (every
        per-table structure is built again), the job, its results."""
        tl, tr = ctx.from_device(dl), ctx.from_device(dr)
        return ctx.submit(dp, tl, tr), (tl, tr)

    def run_steps(make, steps, inflight):
        """`steps` jobs, at most `inflight` of them submitted and not yet
        finished.  A producer thread submits (tables and job) while this
        thread collects results in submission order, as the reference's
        producer and reducer threads do (host_producer.h, host_reducer.h);
        the library's calls release the interpreter lock."""
        stats, last = [], None
        ready = queue.Queue()
        slots = threading.Semaphore(max(1, inflight))
        failure, stop = [], threading.Event()

        def produce():
            try:
                for _ in range(steps):
                    slots.acquire()
                    if stop.is_set():
                        return
                    ready.put(make())
            except BaseException as e:  # handed to the collecting thread
                failure.append(e)
                ready.put(None)

        t0 = time.perf_counter()
        producer = threading.Thread(target=produce)
        producer.start()
        try:
            for _ in range(steps):
                item = ready.get()
                if item is None:
                    raise failure[0]
                last, st = finish(item)
                stats.append(st)
                slots.release()
        finally:
            stop.set()
            slots.release()
            producer.join()
        ctx.synchronize()
        return time.perf_counter() - t0, stats, last

    t = time.perf_counter()
    pairs, cold = finish(fresh())
    cold_s = time.perf_counter() - t
    run_steps(fresh, max(0, args.warmup - 1), args.inflight)
    dt, stats, pairs = run_steps(fresh, args.steps, args.inflight)
    ms_per_step = dt * 1e3 / args.steps
    value = n_pairs * args.steps / dt

    # PCIe-inclusive: the same steps with the rows handed over from host
    # memory (hb_table_upload_serialized: one copy in, unpacked on the device).
    dt_pcie, _, _ = run_steps(lambda: (lambda tl, tr: (ctx.submit(dp, tl, tr), (tl, tr)))(
        ctx.upload_serialized(sl), ctx.upload_serialized(sr)), args.steps, args.inflight)
    # Resident and warm: one pair of tables kept across jobs, so per-table
    # structures are built once (the repeated-job rate of a resident table).
    tl, tr = ctx.from_device(dl), ctx.from_device(dr)
    for _ in range(2):
        finish((ctx.submit(dp, tl, tr), ()))
    dt_warm, stats_warm, _ = run_steps(lambda: (ctx.submit(dp, tl, tr), ()),
                                       max(args.steps, 20), args.inflight)
    # Jobs alone, one at a time, on the warm tables: their latency, and the
    # kernels' durations undisturbed (jobs in flight overlap on the GPU and
    # stretch each other's kernels), so the roofline is taken from these.
    lat, alone = [], []
    for _ in range(max(5, args.steps // 2)):
        t = time.perf_counter()
        _, st_alone = finish((ctx.submit(dp, tl, tr), ()))
        lat.append((time.perf_counter() - t) * 1e3)
        alone.append(st_alone)
    # A fresh job alone: its latency from the device rows to the results.
    lat_fresh, fresh_dev = [], []
    for _ in range(5):
        t = time.perf_counter()
        _, st = finish(fresh())
        lat_fresh.append((time.perf_counter() - t) * 1e3)
        fresh_dev.append(st["device_ms"])

    filt_ms = float(np.mean([s["filter_ms"] for s in alone]))
    filt_ms_timed = float(np.mean([s["filter_ms"] for s in stats]))
    cand = stats[-1]["candidates"]
    achieved = cand * FILTER_OPS_PER_CANDIDATE / (filt_ms * 1e-3) / 1e12
    traffic = measured_traffic(args.workload)
    roofline = {"bound": "valu", "kernel": "k_filter_tiles",
                "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tops/s",
                "frac": achieved / VALU_PEAK_TOPS, "traffic": traffic,
                "ops_per_candidate": FILTER_OPS_PER_CANDIDATE, "candidates_per_launch": cand,
                "avg_launch_ms": filt_ms,
                "launch_ms_from": "jobs run alone on warm tables (HIP events on the job's "
                                  "stream); with %d fresh jobs in flight the filter stage took "
                                  "%.4f ms" % (args.inflight, filt_ms_timed),
                "measured_valu_instr_per_candidate": measured_valu_per_candidate(args.workload),
                "hbm_gbs": None if traffic is None else traffic / (filt_ms * 1e-3) / 1e9,
                "hbm_peak_gbs": HBM_PEAK_GBS}
    served = None
    if not args.no_serve:
        from hyperblocker_amd import workloads
        served = workloads.run_c5(ctx, n_jobs=args.serve_jobs,
                                  log=lambda m: print(m, file=sys.stderr, flush=True))
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
                               "JW author>=0.85), n_partitions=1; each step a fresh job from "
                               "the reference's row-major tables resident in HBM (per-table "
                               "builds, kernels and the copy of the pairs to the host inside "
                               "the step)" % (args.workload, cfg["n_l"], cfg["n_r"]),
                   "pairs_per_step": n_pairs, "matches": int(len(pairs)),
                   "inflight": args.inflight},
        "fresh_job_alone_ms": float(np.median(lat_fresh)),
        "fresh_job_alone_device_ms": float(np.median(fresh_dev)),
        "pcie_inclusive_pairs_per_s": n_pairs * args.steps / dt_pcie,
        "pcie_inclusive_ms_per_step": dt_pcie * 1e3 / args.steps,
        "resident_warm_pairs_per_s": n_pairs * max(args.steps, 20) / dt_warm,
        "resident_warm_ms_per_step": dt_warm * 1e3 / max(args.steps, 20),
        "cold_first_job_ms": cold_s * 1e3,
        # per job on warm tables, alone; fresh jobs in the timed loop (overlapping)
        "stage_ms": {k: float(np.mean([s[k] for s in alone]))
                     for k in ("device_ms", "join_ms", "filter_ms", "exact_ms")},
        "stage_ms_timed": {k: float(np.mean([s[k] for s in stats]))
                           for k in ("device_ms", "join_ms", "filter_ms", "exact_ms")},
        "job_alone_ms": float(np.median(lat)),
        "filtered": stats[-1]["filtered"], "jw_exact": stats[-1]["jw_exact"],
        "roofline": roofline, "cpu_baseline": cpu,
        # c5: served (continuous batching) against plain launches on one trace
        "served": None if served is None else {
            k: served[k] for k in ("jobs", "right_rows", "arrival_rate_per_s", "offered_load",
                                   "plain_service_ms", "served_pairs_per_s", "served_p50_ms",
                                   "served_p99_ms", "plain_pairs_per_s", "plain_p50_ms",
                                   "plain_p99_ms", "saturated_served_pairs_per_s",
                                   "saturated_served_p50_ms", "saturated_served_p99_ms",
                                   "saturated_plain_pairs_per_s", "saturated_plain_p50_ms",
                                   "saturated_plain_p99_ms", "served_batches")},
        # run_hyperblocker_exec on the same tables as CSV files (fresh process)
        "end_to_end_cli": e2e,
    }
    print(json.dumps(out))
    dp.release()
    tl.release()
    tr.release()
    dl.free()
    dr.free()
    ctx.close()


if __name__ == "__main__":
    main()
