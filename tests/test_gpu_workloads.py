"""The batched (c4) and served (c5) job mixes on the HIP path, scaled to
run in about a minute (needs an MI355X).

c4: 128 jobs drawn as in workloads.c4_specs (per-side rows log-uniform on
[100, 50k], 1-4 rules from the streamed workload's templates), as slices of
two resident 200,000-row pools of the 8-column schema.  c5: 100 left jobs
of 1-10k rows (workloads.c5_specs, 1-3 rules) against a resident
200,000-row right table, arriving at random intervals at a serving loop.

Every job's result in the batch, or served in the mix, equals the same job
run alone (pairs and audit entries), and every 16th job equals the CPU
oracle on every k-th left row (k chosen so that the oracle evaluates at
most 2e7 pairs per job).  Reference counterpart: the per-partition launch
loop, core/components/host_producer.h:75-104.
"""
import math
import random
import threading
import time

import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import datagen, workloads
from tests import helpers as H

pytestmark = pytest.mark.gpu

ORACLE_PAIRS = 2e7


@pytest.fixture(scope="module")
def ctx():
    c = hb.Context(0)
    yield c
    c.close()


def _plan(rules):
    return hb.HostPlan.from_yaml([workloads.TEMPLATES[r] for r in rules])


def _results(job):
    pairs, audit, stats = job.results()
    job.release()
    return hb.sort_pairs(pairs), H.sort_audit(audit), stats


def _check_sampled(plan, L, R, got, what):
    """The job's pairs and audit entries on every k-th left row equal the
    oracle's there."""
    n_l, n_r = len(L[0]), len(R[0])
    step = max(1, math.ceil(n_l * n_r / ORACLE_PAIRS))
    want, want_audit, _ = H.run_oracle(plan, L, R, l_step=step)
    pairs, audit = got[0], got[1]
    H.assert_same_pairs(pairs[pairs[:, 0] % step == 0], want, what + " (every %d-th row)" % step)
    H.assert_same_audit(audit[audit["l"] % step == 0], want_audit, what + " audit")


def test_c4_job_mix_batched_equals_alone_and_oracle(ctx):
    L, R, _ = datagen.generate(4, 200_000, 200_000, 170_000, datagen.SCHEMA8)
    specs = workloads.c4_specs(n_jobs=128, pool=200_000)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    host_plans, plans, slices, jobs = {}, {}, [], []
    try:
        for off_l, n_l, off_r, n_r, rules in specs:
            if rules not in plans:
                host_plans[rules] = _plan(rules)
                plans[rules] = ctx.plan(host_plans[rules])
            sl, sr = tl.slice(off_l, off_l + n_l), tr.slice(off_r, off_r + n_r)
            slices += [sl, sr]
            jobs.append((plans[rules], sl, sr))
        batched = [_results(j) for j in ctx.submit_batch(jobs)]
        for k, ((p, sl, sr), spec, b) in enumerate(zip(jobs, specs, batched)):
            off_l, n_l, off_r, n_r, rules = spec
            what = "c4 job %d (%d x %d, rules %s)" % (k, n_l, n_r, rules)
            alone = _results(ctx.submit(p, sl, sr))
            H.assert_same_pairs(b[0], alone[0], what + " batch vs alone")
            H.assert_same_audit(b[1], alone[1], what + " batch vs alone")
            assert b[2]["pairs"] == n_l * n_r
            if k % 16 == 0:
                _check_sampled(host_plans[rules], [c[off_l:off_l + n_l] for c in L],
                               [c[off_r:off_r + n_r] for c in R], b, what)
    finally:
        for s in slices:
            s.release()
        tl.release()
        tr.release()
        for p in plans.values():
            p.release()


def test_c5_job_mix_served_equals_alone_and_oracle(ctx):
    L, R, _ = datagen.generate(5, 50_000, 200_000, 40_000, datagen.SCHEMA8)
    specs = workloads.c5_specs(100, pool=50_000)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    host_plans, plans = {}, {}
    slices = [tl.slice(off, off + n) for off, n, _ in specs]
    try:
        for _, _, rules in specs:
            if rules not in plans:
                host_plans[rules] = _plan(rules)
                plans[rules] = ctx.plan(host_plans[rules])
        rng = random.Random(55)
        served = [None] * len(specs)
        with hb.Server(ctx, max_batches_in_flight=2, max_jobs_per_batch=16) as srv:
            handles = []
            for (_, _, rules), sl in zip(specs, slices):
                handles.append(srv.submit(plans[rules], sl, tr))
                time.sleep(rng.expovariate(2000.0))  # about 0.5 ms between arrivals
            for k, h in enumerate(handles):
                served[k] = _results(h)
            info = srv.info()
        assert info["jobs"] == len(specs) and info["batches"] < len(specs)
        for k, ((off, n, rules), sl, s) in enumerate(zip(specs, slices, served)):
            what = "c5 job %d (%d x 200000, rules %s)" % (k, n, rules)
            alone = _results(ctx.submit(plans[rules], sl, tr))
            H.assert_same_pairs(s[0], alone[0], what + " served vs alone")
            H.assert_same_audit(s[1], alone[1], what + " served vs alone")
            assert s[2]["latency_ms"] >= s[2]["queue_ms"] >= 0
            if k % 16 == 0:
                _check_sampled(host_plans[rules], [c[off:off + n] for c in L], R, s, what)
    finally:
        for s in slices:
            s.release()
        tl.release()
        tr.release()
        for p in plans.values():
            p.release()
