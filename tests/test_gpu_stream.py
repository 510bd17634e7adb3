"""Streamed partitions on the HIP path (needs an MI355X).

Partitions of a three-rule, eight-column workload are uploaded, run and
released while other partitions are in flight, twice over (the second pass
reuses the memory the first one released).  Every partition's pairs and
audit entries equal the oracle's on that partition's rows.
"""
import collections

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


def _stream(ctx, dp, lt, rt, parts, inflight):
    got = {}
    pending = collections.deque()

    def drain():
        k, job, tl, tr = pending.popleft()
        pairs, audit, _ = job.results()
        job.release()
        tl.release()
        tr.release()
        got[k] = (hb.sort_pairs(pairs), H.sort_audit(audit))

    for k, (lr, rr) in enumerate(parts):
        tl, tr = ctx.upload(lt, rows=lr), ctx.upload(rt, rows=rr)
        pending.append((k, ctx.submit(dp, tl, tr), tl, tr))
        if len(pending) >= inflight:
            drain()
    while pending:
        drain()
    return got


@pytest.mark.parametrize("n_partitions,inflight", [(16, 3), (5, 1)])
def test_streamed_partitions_equal_oracle(ctx, n_partitions, inflight):
    L, R, _, cfg = datagen.make("c3", scale=(6000, 5000))
    lt, rt = hb.ColumnTable(L), hb.ColumnTable(R)
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    parts = hb.partition(lt, rt, plan, n_partitions)
    assert 1 < len(parts) <= n_partitions
    dp = ctx.plan(plan)
    try:
        first = _stream(ctx, dp, lt, rt, parts, inflight)
        second = _stream(ctx, dp, lt, rt, parts, inflight)
        for k, (lr, rr) in enumerate(parts):
            want, want_audit, _ = H.run_oracle(plan, [c[lr] for c in L], [c[rr] for c in R])
            for got in (first, second):
                H.assert_same_pairs(got[k][0], want, "partition %d" % k)
                H.assert_same_audit(got[k][1], want_audit, "partition %d" % k)
    finally:
        dp.release()
