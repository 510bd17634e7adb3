"""Served form and table slices on the HIP path (needs an MI355X).

A server keeps jobs in a queue and admits them into batches while others
run (continuous batching).  Every job's result must be the same alone and
in the mix; slices of a resident table must give the same results as the
same rows uploaded as a table of their own.
"""
import random
import threading
import time

import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import datagen
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hb.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def served(ctx):
    """A resident right table (c1 right side) and a resident left pool, with
    left jobs as slices of the pool."""
    L, R, dup, cfg = datagen.make("c1")
    rules = datagen.RULES_C3[:1] + [datagen.rule_yaml([3], [(1, 0.8)])]
    plans = [ctx.plan(hb.HostPlan.from_yaml(rules[:1])), ctx.plan(hb.HostPlan.from_yaml(rules))]
    tl = ctx.upload(hb.ColumnTable(L))
    tr = ctx.upload(hb.ColumnTable(R))
    yield L, R, plans, tl, tr
    tl.release()
    tr.release()
    for p in plans:
        p.release()


def _finish(job):
    pairs, audit, stats = job.results()
    job.release()
    return hb.sort_pairs(pairs), H.sort_audit(audit), stats


def _specs(n, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        size = rng.choice([1, 10, 90, 400, 1200])
        begin = rng.randrange(0, 2600 - size)
        out.append((begin, begin + size, rng.randrange(2)))
    return out


def test_slices_equal_uploaded_rows(ctx, served):
    L, R, plans, tl, tr = served
    for begin, end, pk in _specs(6, 1):
        sl = tl.slice(begin, end)
        own = ctx.upload(hb.ColumnTable([c[begin:end] for c in L]))
        try:
            a = _finish(ctx.submit(plans[pk], sl, tr))
            b = _finish(ctx.submit(plans[pk], own, tr))
            H.assert_same_pairs(a[0], b[0], "slice [%d, %d)" % (begin, end))
            H.assert_same_audit(a[1], b[1], "slice [%d, %d)" % (begin, end))
        finally:
            sl.release()
            own.release()
    with pytest.raises(hb.HBError):
        tl.slice(10, 5)
    sl = tl.slice(0, 10)
    with pytest.raises(hb.HBError):  # the parent cannot go while a slice lives
        tl.release()
    sl.release()


@pytest.mark.parametrize("budget", [0, 1])
def test_served_results_equal_alone(ctx, served, budget):
    """60 jobs arrive from another thread at random intervals; the server
    batches them as slots free up (or one at a time when the memory budget
    admits only one).  Each job's pairs and audit list equal the same job run
    alone, and the first few equal the oracle."""
    L, R, plans, tl, tr = served
    specs = _specs(60, 7 + budget)
    slices = [tl.slice(b, e) for b, e, _ in specs]
    try:
        alone = [_finish(ctx.submit(plans[pk], sl, tr)) for (_, _, pk), sl in zip(specs, slices)]
        jobs = [None] * len(specs)
        with hb.Server(ctx, max_batches_in_flight=2, max_jobs_per_batch=16,
                       memory_budget=budget) as srv:
            def producer():
                rng = random.Random(3)
                for k, ((_, _, pk), sl) in enumerate(zip(specs, slices)):
                    jobs[k] = srv.submit(plans[pk], sl, tr)
                    time.sleep(rng.choice([0, 0, 0.0002, 0.001]))
            th = threading.Thread(target=producer)
            th.start()
            th.join()
            mixed = [_finish(j) for j in jobs]
            info = srv.info()
        assert info["jobs"] == len(specs)
        if budget == 1:
            assert info["batches"] == len(specs)  # memory admits one job at a time
        for k, (a, m) in enumerate(zip(alone, mixed)):
            H.assert_same_pairs(m[0], a[0], "job %d" % k)
            H.assert_same_audit(m[1], a[1], "job %d" % k)
            assert m[2]["latency_ms"] >= m[2]["queue_ms"] >= 0
        for k in range(3):
            b, e, pk = specs[k]
            plan = hb.HostPlan.from_yaml((datagen.RULES_C3[:1] +
                                          [datagen.rule_yaml([3], [(1, 0.8)])])[:pk + 1])
            want, want_audit, _ = H.run_oracle(plan, [c[b:e] for c in L], R)
            H.assert_same_pairs(mixed[k][0], want, "job %d vs oracle" % k)
    finally:
        for sl in slices:
            sl.release()


def test_server_batches_concurrent_arrivals(ctx, served):
    """Jobs submitted in a burst while a large job runs are admitted together
    into few batches."""
    L, R, plans, tl, tr = served
    slices = [tl.slice(k * 10, k * 10 + 10) for k in range(40)]
    try:
        with hb.Server(ctx, max_batches_in_flight=1, max_jobs_per_batch=64) as srv:
            first = srv.submit(plans[1], tl, tr)
            burst = [srv.submit(plans[0], sl, tr) for sl in slices]
            for j in [first] + burst:
                j.wait()
            info = srv.info()
        assert info["jobs"] == 41 and info["batches"] < 41
        _finish(first)
        for j in burst:
            _finish(j)
    finally:
        for sl in slices:
            sl.release()
