"""Batched and served blocking workloads (the job-set forms of the path).

Batched (c4): 4,096 independent jobs, per-side table sizes log-uniform on
[100, 50k] rows, each job using 1-4 rules drawn from four templates (the
three rules of the streamed 1M x 1M workload plus venue equality with
JW(city) >= 0.9), 8-column tables of the streamed workload's schema.  Job
tables are slices of two resident 1M-row pools.  The batch runs as one
launch sequence (hb_batch_submit) and is compared with the same 4,096 jobs
submitted one by one (hb_job_submit, a few in flight).

Served (c5): a 2M-row right table resident in HBM; left jobs of 1-10k rows
(slices of a resident 200k-row pool) carrying 1-3 rules arrive as a Poisson
stream at 70% of the measured capacity of plain sequential jobs.  They are
served by the continuous-batching server (hb_server_*) and, on the same
arrival trace, by plain one-at-a-time launches; both report p50/p99 latency
(arrival to results ready) and throughput.

Streamed (c3): 1M x 1M rows of the 8-column schema, three rules, split
into 16 partitions by the reference's hash partitioning; each partition's
rows are packed and copied to the device while earlier partitions compute
(a few partitions in flight).  Pairs/s counts every stage after
partitioning: host packing, host-to-device copies, kernels and result copies.

Nothing here computes a predicate: every job runs through the C ABI.
"""
import collections
import concurrent.futures
import math
import random
import threading
import time

import numpy as np

from . import blocker as hbk
from . import datagen

TEMPLATES = datagen.RULES_C3 + [datagen.rule_yaml([3], [(5, 0.9)])]


def c4_specs(n_jobs=4096, seed=4, pool=1_000_000, lo=100, hi=50_000):
    """[(off_l, n_l, off_r, n_r, rule indices)] of the batched workload."""
    rng = random.Random(seed)
    out = []
    for _ in range(n_jobs):
        n_l = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        n_r = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        k = rng.randint(1, 4)
        rules = tuple(sorted(rng.sample(range(len(TEMPLATES)), k)))
        out.append((rng.randrange(0, pool - n_l), n_l, rng.randrange(0, pool - n_r), n_r, rules))
    return out


def c5_specs(n_jobs, seed=5, pool=200_000, lo=1, hi=10_000):
    """[(off_l, n_l, rule indices)] of the served workload (1-3 rules)."""
    rng = random.Random(seed)
    out = []
    for _ in range(n_jobs):
        n_l = rng.randint(lo, hi)
        k = rng.randint(1, 3)
        rules = tuple(sorted(rng.sample(range(len(TEMPLATES)), k)))
        out.append((rng.randrange(0, pool - n_l), n_l, rules))
    return out


class _Plans:
    """Device plans per rule subset, created on first use."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.plans = {}

    def __getitem__(self, rules):
        if rules not in self.plans:
            self.plans[rules] = self.ctx.plan(
                hbk.HostPlan.from_yaml([TEMPLATES[r] for r in rules]))
        return self.plans[rules]

    def release(self):
        for p in self.plans.values():
            p.release()


def _drain(jobs):
    """Results of finished jobs copied out, jobs released; returns pairs."""
    n = 0
    for j in jobs:
        pairs, _, _ = j.results()
        n += len(pairs)
        j.release()
    return n


def run_c4(ctx, n_jobs=4096, reps=2, inflight=4, autotune=True, log=print):
    t0 = time.perf_counter()
    L, R, _ = datagen.generate(4, 1_000_000, 1_000_000, 850_000, datagen.SCHEMA8)
    tl = ctx.upload(hbk.ColumnTable(L))
    tr = ctx.upload(hbk.ColumnTable(R))
    ctx.synchronize()
    log("c4: pools generated and uploaded in %.1f s" % (time.perf_counter() - t0))
    specs = c4_specs(n_jobs)
    plans = _Plans(ctx)
    slices = []
    jobs = []
    for off_l, n_l, off_r, n_r, rules in specs:
        sl, sr = tl.slice(off_l, off_l + n_l), tr.slice(off_r, off_r + n_r)
        slices += [sl, sr]
        jobs.append((plans[rules], sl, sr))
    pairs_total = sum(n_l * n_r for _, n_l, _, n_r, _ in specs)
    try:
        t = time.perf_counter()
        _drain(ctx.submit_batch(jobs))  # per-table structures (layouts) built here
        cold_s = time.perf_counter() - t
        log("c4: cold batch %.2f s" % cold_s)
        tuned = None
        if autotune:
            t = time.perf_counter()
            sample = jobs[:: max(1, len(jobs) // 512)]
            cands = [(1, 256, 16), (2, 128, 32), (2, 192, 16), (2, 192, 32), (2, 256, 16), (2, 256, 32),
                     (2, 512, 16), (4, 256, 12), (4, 512, 8)]
            tuned = ctx.autotune(sample, cands)
            log("c4: autotuned shape %s (filter %.3f ms on %d jobs) in %.1f s"
                % (tuned[0], tuned[1], len(sample), time.perf_counter() - t))
        # batched: one launch sequence for all jobs
        batched = []
        stats = None
        for _ in range(reps):
            t = time.perf_counter()
            handles = ctx.submit_batch(jobs)
            _, pairs, _, _, _ = ctx.batch_results(handles[0])  # every job's results
            stats = handles[0].stats()
            matches = len(pairs)
            ctx.release_jobs(handles)
            batched.append(time.perf_counter() - t)
        # captured: the same batch as one HIP graph, replayed
        graph = ctx.batch_graph(jobs)
        try:
            graphed = []
            gstats = None
            for _ in range(reps + 1):
                t = time.perf_counter()
                handles = graph.run()
                _, gpairs, _, _, _ = ctx.batch_results(handles[0])
                gstats = handles[0].stats()
                ctx.release_jobs(handles)
                graphed.append(time.perf_counter() - t)
                if len(gpairs) != matches:
                    raise AssertionError("graph run: %d pairs, batch: %d" % (len(gpairs), matches))
        finally:
            graph.release()
        # plain: one job per submit, a few in flight
        plain = []
        for _ in range(reps):
            t = time.perf_counter()
            pending = []
            for p, sl, sr in jobs:
                pending.append(ctx.submit(p, sl, sr))
                if len(pending) >= inflight:
                    _drain([pending.pop(0)])
            _drain(pending)
            plain.append(time.perf_counter() - t)
        bt, pt, gt = min(batched), min(plain), min(graphed[1:])
        return {
            "jobs": n_jobs, "pairs": pairs_total, "matches": matches,
            "batched_s": bt, "plain_s": pt, "graph_s": gt,
            "batched_pairs_per_s": pairs_total / bt, "plain_pairs_per_s": pairs_total / pt,
            "graph_pairs_per_s": pairs_total / gt, "graph_first_run_s": graphed[0],
            "graph_device_ms": gstats["device_ms"],
            "speedup": pt / bt, "cold_s": cold_s,
            "batch_device_ms": stats["device_ms"], "batch_filter_ms": stats["filter_ms"],
            "batch_exact_ms": stats["exact_ms"], "batch_join_ms": stats["join_ms"],
            "filter_shape": list(tuned[0]) if tuned else list(ctx.filter_shape),
        }
    finally:
        for s in slices:
            s.release()
        tl.release()
        tr.release()
        plans.release()


def _percentiles(xs):
    a = np.sort(np.asarray(xs))
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


def _serve_trace(submit, specs, arrivals, plans, tables, n_l_pairs):
    """Submits specs[k] at arrivals[k] seconds (busy-wait clock) and collects
    each job's stats.  Returns (latencies ms, wall s)."""
    handles = [None] * len(specs)
    lat = []
    errors = []
    stop = threading.Event()

    def consume():
        try:
            for k in range(len(specs)):
                while handles[k] is None:
                    if stop.is_set():
                        return
                    time.sleep(0.0001)
                j = handles[k]
                j.wait()
                lat.append(j.stats()["latency_ms"])
                j.release()
        except Exception as e:  # reported by the submitting thread
            errors.append(RuntimeError("job %d of %d: %s" % (len(lat), len(specs), e)))

    th = threading.Thread(target=consume)
    th.start()
    t0 = time.perf_counter()
    try:
        for k, ((off_l, n_l, rules), at) in enumerate(zip(specs, arrivals)):
            while True:  # sleep (letting the consumer run), then spin the last 0.3 ms
                left = at - (time.perf_counter() - t0)
                if left <= 0:
                    break
                if left > 0.0006:
                    time.sleep(left - 0.0003)
            if errors:
                break
            handles[k] = submit(plans[rules], tables[k])
    finally:
        stop.set()
        th.join()
    if errors:
        raise errors[0]
    return lat, time.perf_counter() - t0


def _mem(ctx):
    m = ctx.memory()
    return "device free %.1f of %.1f GiB, library live %.1f GiB (cached %.1f)" % (
        m["device_free"] / 2**30, m["device_total"] / 2**30, m["pool_live"] / 2**30,
        m["pool_cached"] / 2**30)


def run_c5(ctx, n_jobs=1500, load=0.7, log=print):
    t0 = time.perf_counter()
    L, R, _ = datagen.generate(5, 200_000, 2_000_000, 170_000, datagen.SCHEMA8)
    tl = ctx.upload(hbk.ColumnTable(L))
    tr = ctx.upload(hbk.ColumnTable(R))
    ctx.synchronize()
    log("c5: tables generated and uploaded in %.1f s" % (time.perf_counter() - t0))
    specs = c5_specs(n_jobs)
    plans = _Plans(ctx)
    slices = [tl.slice(off, off + n) for off, n, _ in specs]
    pairs_total = sum(n * 2_000_000 for _, n, _ in specs)
    try:
        # warm: per-table structures of every slice, and plain capacity
        t = time.perf_counter()
        _drain(ctx.submit_batch([(plans[r], sl, tr) for (_, _, r), sl in zip(specs, slices)]))
        log("c5: cold batch %.2f s; %s" % (time.perf_counter() - t, _mem(ctx)))
        n_cap = min(200, n_jobs)
        t = time.perf_counter()
        for (_, _, r), sl in zip(specs[:n_cap], slices[:n_cap]):
            _drain([ctx.submit(plans[r], sl, tr)])
        service = (time.perf_counter() - t) / n_cap
        lam = load / service
        rng = random.Random(55)
        arrivals = []
        at = 0.0
        for _ in range(n_jobs):
            at += rng.expovariate(lam)
            arrivals.append(at)
        log("c5: plain service %.3f ms/job, arrival rate %.0f jobs/s; %s"
            % (service * 1e3, lam, _mem(ctx)))
        # served: continuous batching
        with hbk.Server(ctx, max_batches_in_flight=2, max_jobs_per_batch=64) as srv:
            lat_s, wall_s = _serve_trace(lambda p, sl: srv.submit(p, sl, tr), specs, arrivals,
                                         plans, slices, pairs_total)
            info = srv.info()
        log("c5: served trace done; %s" % _mem(ctx))
        # plain: the same trace, one job at a time
        with hbk.Server(ctx, max_batches_in_flight=1, max_jobs_per_batch=1) as srv:
            lat_p, wall_p = _serve_trace(lambda p, sl: srv.submit(p, sl, tr), specs, arrivals,
                                         plans, slices, pairs_total)
        # saturated: every job queued at once (throughput when the offered
        # load exceeds what plain launches sustain)
        now = [0.0] * n_jobs
        with hbk.Server(ctx, max_batches_in_flight=2, max_jobs_per_batch=64) as srv:
            lat_ss, wall_ss = _serve_trace(lambda p, sl: srv.submit(p, sl, tr), specs, now,
                                           plans, slices, pairs_total)
        with hbk.Server(ctx, max_batches_in_flight=1, max_jobs_per_batch=1) as srv:
            lat_sp, wall_sp = _serve_trace(lambda p, sl: srv.submit(p, sl, tr), specs, now,
                                           plans, slices, pairs_total)
        log("c5: saturated: served %.3f s, plain %.3f s" % (wall_ss, wall_sp))
        p50s, p99s = _percentiles(lat_s)
        p50p, p99p = _percentiles(lat_p)
        return {
            "jobs": n_jobs, "right_rows": 2_000_000, "pairs": pairs_total,
            "arrival_rate_per_s": lam, "offered_load": load,
            "plain_service_ms": service * 1e3,
            "served_pairs_per_s": pairs_total / wall_s, "served_p50_ms": p50s,
            "served_p99_ms": p99s, "served_batches": info["batches"],
            "plain_pairs_per_s": pairs_total / wall_p, "plain_p50_ms": p50p,
            "plain_p99_ms": p99p,
            "saturated_served_pairs_per_s": pairs_total / wall_ss,
            "saturated_served_p50_ms": _percentiles(lat_ss)[0],
            "saturated_served_p99_ms": _percentiles(lat_ss)[1],
            "saturated_plain_pairs_per_s": pairs_total / wall_sp,
            "saturated_plain_p50_ms": _percentiles(lat_sp)[0],
            "saturated_plain_p99_ms": _percentiles(lat_sp)[1],
        }
    finally:
        for s in slices:
            s.release()
        tl.release()
        tr.release()
        plans.release()


def run_c3(ctx, inflight=3, reps=2, log=print):
    """c3 streamed: each hash partition's rows are uploaded while earlier
    partitions run.  Two handoffs are timed: column arrays gathered and
    packed by the host (hb_table_upload_rows), and the reference's
    row-major SerializedTable, one row copy per partition row with the cells
    unpacked on the device (hb_table_upload_serialized_rows)."""
    t = time.perf_counter()
    L, R, _, cfg = datagen.make("c3")
    lt, rt = hbk.ColumnTable(L), hbk.ColumnTable(R)
    plan = hbk.HostPlan.from_yaml(cfg["rules"])
    log("c3: tables generated in %.1f s" % (time.perf_counter() - t))
    t = time.perf_counter()
    parts = hbk.partition(lt, rt, plan, cfg["n_partitions"])
    part_s = time.perf_counter() - t
    pairs_total = sum(len(a) * len(b) for a, b in parts)
    log("c3: %d partitions (%.3g pairs) in %.2f s" % (len(parts), pairs_total, part_s))
    sl, sr = hbk.SerializedTable(L), hbk.SerializedTable(R)
    dp = ctx.plan(plan)
    uploads = {
        "columns": lambda lr, rr: (ctx.upload(lt, rows=lr), ctx.upload(rt, rows=rr)),
        "serialized": lambda lr, rr: (ctx.upload_serialized(sl, rows=lr),
                                      ctx.upload_serialized(sr, rows=rr)),
    }

    def stream(upload):
        """Partitions in order: a producer thread uploads up to two
        partitions ahead (host_producer.h's role) while this thread submits
        and drains; upload_s is the producer's time, overlapped."""
        t0 = time.perf_counter()
        pending = collections.deque()
        out = {"matches": 0, "device_ms": 0.0, "upload_s": 0.0, "submit_s": 0.0,
               "drain_s": 0.0, "candidates": 0, "filtered": 0, "jw_exact": 0, "audit": 0}

        def produce(lr, rr):
            tu = time.perf_counter()
            tables = upload(lr, rr)
            return tables, time.perf_counter() - tu

        def drain():
            td = time.perf_counter()
            job, tl, tr = pending.popleft()
            pairs, audit, st = job.results()
            job.release()
            tl.release()
            tr.release()
            out["matches"] += len(pairs)
            out["audit"] += len(audit)
            out["device_ms"] += st["device_ms"]
            out["candidates"] += st["candidates"]
            out["filtered"] += st["filtered"]
            out["jw_exact"] += st["jw_exact"]
            out["drain_s"] += time.perf_counter() - td

        with concurrent.futures.ThreadPoolExecutor(1) as producer:
            ahead = collections.deque(producer.submit(produce, lr, rr) for lr, rr in parts[:2])
            nxt = 2
            while ahead:
                (tl, tr), du = ahead.popleft().result()
                out["upload_s"] += du
                if nxt < len(parts):
                    ahead.append(producer.submit(produce, *parts[nxt]))
                    nxt += 1
                ts = time.perf_counter()
                pending.append((ctx.submit(dp, tl, tr), tl, tr))
                out["submit_s"] += time.perf_counter() - ts
                if len(pending) >= inflight:
                    drain()
        while pending:
            drain()
        out["wall_s"] = time.perf_counter() - t0
        return out

    try:
        out = {"partitions": len(parts), "pairs": pairs_total, "partition_s": part_s,
               "inflight": inflight}
        for handoff, upload in uploads.items():
            runs = []
            for k in range(reps + 1):
                r = stream(upload)
                log("c3 (%s): stream %d: %.3f s (host: uploads %.3f s, submits %.3f s, "
                    "drains %.3f s; device %.1f ms)" % (handoff, k, r["wall_s"], r["upload_s"],
                                                       r["submit_s"], r["drain_s"],
                                                       r["device_ms"]))
                runs.append(r)
            best = min(runs[1:], key=lambda r: r["wall_s"])
            res = {
                "streamed_s": best["wall_s"], "streamed_pairs_per_s": pairs_total / best["wall_s"],
                "first_stream_s": runs[0]["wall_s"], "device_ms": best["device_ms"],
                "device_pairs_per_s": pairs_total / (best["device_ms"] * 1e-3),
                "host_upload_s": best["upload_s"], "host_submit_s": best["submit_s"],
                "host_drain_s": best["drain_s"], "matches": best["matches"],
                "candidates": best["candidates"], "filtered": best["filtered"],
                "jw_exact": best["jw_exact"],
                "audit": best["audit"],
            }
            if handoff == "columns":
                out.update(res)
            else:
                out.update({"serialized_" + k: v for k, v in res.items()})
        if out["serialized_matches"] != out["matches"] or \
                out["serialized_audit"] != out["audit"]:
            raise RuntimeError("c3: the two handoffs gave different results")
        return out
    finally:
        dp.release()
