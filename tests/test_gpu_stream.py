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


def test_c3_full_size_sampled_rows_and_pair_properties(ctx):
    """The streamed workload at full size (1M x 1M rows, 8 columns, three
    rules, 16 hash partitions, 3 in flight):
      * every 100th left row of each partition: pairs equal the oracle's on
        those rows, and so do audit entries on every 500th left row;
      * every emitted pair satisfies the plan (the oracle's per-pair
        evaluation of all of them);
      * every generator duplicate (right row k copies left row dup[k]) whose
        two rows share a partition and that satisfies the plan is emitted.
    """
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from oracle import oracle as O

    L, R, dup, cfg = datagen.make("c3")
    lt, rt = hb.ColumnTable(L), hb.ColumnTable(R)
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    parts = hb.partition(lt, rt, plan, cfg["n_partitions"])
    dp = ctx.plan(plan)
    try:
        got = _stream(ctx, dp, lt, rt, parts, 3)
    finally:
        dp.release()
    print("c3: %d partitions streamed, %d pairs" % (len(parts), sum(len(g[0]) for g in got.values())),
          flush=True)
    op = H.oracle_plan(plan)
    tl_all, tr_all = O.Table(L), O.Table(R)

    def sampled(k):
        lr, rr = parts[k]
        Lp, Rp = [c[lr] for c in L], [c[rr] for c in R]
        pairs, _ = O.block(op, O.Table(Lp), O.Table(Rp), l_step=100, n_threads=4)
        audit = O.audit(op, O.Table(Lp), O.Table(Rp), l_step=500, n_threads=4)
        return k, pairs, audit

    with ThreadPoolExecutor(max_workers=4) as ex:
        for k, want, want_audit in ex.map(sampled, range(len(parts))):
            pairs, audit = got[k]
            H.assert_same_pairs(pairs[pairs[:, 0] % 100 == 0], want, "partition %d, rows %% 100" % k)
            H.assert_same_audit(audit[audit["l"] % 500 == 0], want_audit, "partition %d audit" % k)
            print("c3: partition %d equals the oracle on its sampled rows" % k, flush=True)
    # all emitted pairs, in global row numbers
    glob = np.concatenate([np.stack([parts[k][0][got[k][0][:, 0]], parts[k][1][got[k][0][:, 1]]], 1)
                           for k in range(len(parts)) if len(got[k][0])]).astype(np.uint32)
    assert O.eval_pairs(op, tl_all, tr_all, glob).all()
    part_l = np.full(len(L[0]), -1, np.int64)
    part_r = np.full(len(R[0]), -2, np.int64)
    for k, (lr, rr) in enumerate(parts):
        part_l[lr] = k
        part_r[rr] = k
    dpairs = np.stack([dup, np.arange(len(dup))], 1).astype(np.uint32)
    same = part_l[dpairs[:, 0]] == part_r[dpairs[:, 1]]
    dpairs = dpairs[same]
    want_dups = dpairs[O.eval_pairs(op, tl_all, tr_all, dpairs)]
    emitted = set(map(tuple, glob.tolist()))
    missing = [tuple(x) for x in want_dups.tolist() if tuple(x) not in emitted]
    assert not missing, missing[:5]
    assert len(want_dups) > 0.5 * len(dup)
