"""Batched jobs on the HIP path (needs an MI355X).

A batch runs many blocking jobs of different table sizes and plans as one
sequence of kernel launches.  Each job's result must equal the same job run
alone and the CPU oracle, pair for pair and audit entry for audit entry.
"""
import random

import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import datagen
from tests import helpers as H
from tests.test_gpu_parity import PLANS, _adversarial_tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hb.Context(0)
    yield c
    c.close()


def _results(job):
    pairs, audit, stats = job.results()
    job.release()
    return hb.sort_pairs(pairs), H.sort_audit(audit), stats


def _jobs(ctx, n, seed):
    """n jobs with tables of 1..900 rows a side and assorted plans."""
    rng = random.Random(seed)
    names = sorted(PLANS)
    out = []
    for k in range(n):
        n_l = rng.choice([0, 1, 7, 60, 300, 900])
        n_r = rng.choice([1, 5, 80, 400, 700])
        alpha = rng.choice([b"ab", b"abcd", b"abcdefgh"])
        L, R = _adversarial_tables(1000 + k, max(n_l, 1), n_r, alpha, rng.choice([8, 30, 60]))
        if n_l == 0:
            L = [c[:0] for c in L]
        name = names[k % len(names)]
        out.append((name, L, R))
    return out


def test_batch_equals_alone_and_oracle(ctx):
    specs = _jobs(ctx, 30, 5)
    plans = {n: ctx.plan(hb.HostPlan.from_yaml(PLANS[n])) for n in PLANS}
    tables = [(ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))) for _, L, R in specs]
    try:
        alone = [_results(ctx.submit(plans[n], tl, tr)) for (n, _, _), (tl, tr) in zip(specs, tables)]
        batch = ctx.submit_batch([(plans[n], tl, tr) for (n, _, _), (tl, tr) in zip(specs, tables)])
        together = [_results(j) for j in batch]
        for k, ((name, L, R), a, t) in enumerate(zip(specs, alone, together)):
            what = "job %d (%s, %d x %d)" % (k, name, len(L[0]), len(R[0]))
            H.assert_same_pairs(t[0], a[0], what + " batch vs alone")
            H.assert_same_audit(t[1], a[1], what + " batch vs alone")
            assert t[2]["pairs"] == a[2]["pairs"] == len(L[0]) * len(R[0])
            assert t[2]["matches"] == len(a[0])
            if k % 3 == 0:
                plan = hb.HostPlan.from_yaml(PLANS[name])
                want, want_audit, _ = H.run_oracle(plan, L, R)
                H.assert_same_pairs(t[0], want, what + " vs oracle")
                H.assert_same_audit(t[1], want_audit, what + " vs oracle")
    finally:
        for tl, tr in tables:
            tl.release()
            tr.release()
        for p in plans.values():
            p.release()


def test_batch_sharing_a_resident_right_table(ctx):
    """Many left jobs against one resident right table (the served shape),
    with multi-rule plans: per-job results equal the oracle's."""
    L_all, R = datagen.make("c1")[:2]
    rules = datagen.RULES_C3[:1] + [datagen.rule_yaml([3], [(1, 0.8)])]
    plan = hb.HostPlan.from_yaml(rules)
    dp = ctx.plan(plan)
    tr = ctx.upload(hb.ColumnTable(R))
    rng = random.Random(3)
    lefts = []
    for k in range(24):
        rows = np.array(sorted(rng.sample(range(2600), rng.choice([1, 13, 200, 700]))))
        lefts.append([c[rows] for c in L_all])
    tls = [ctx.upload(hb.ColumnTable(L)) for L in lefts]
    try:
        jobs = ctx.submit_batch([(dp, tl, tr) for tl in tls])
        for k, (job, L) in enumerate(zip(jobs, lefts)):
            got, audit, stats = _results(job)
            want, want_audit, _ = H.run_oracle(plan, L, R)
            H.assert_same_pairs(got, want, "left job %d" % k)
            H.assert_same_audit(audit, want_audit, "left job %d" % k)
    finally:
        for tl in tls:
            tl.release()
        tr.release()
        dp.release()


def test_batch_regrowth_keeps_every_job_exact(ctx):
    """One job of the batch emits 2.1M pairs (far beyond the first buffers);
    the batch regrows and every job still matches its lone result."""
    rows_l = [(str(i).encode(), b"x", b"y", b"k", b"k") for i in range(1500)]
    rows_r = [(str(i).encode(), b"x", b"y", b"k", b"k") for i in range(1400)]
    big = (H.columns_from_rows(rows_l), H.columns_from_rows(rows_r))
    small = [_adversarial_tables(50 + k, 100, 90, b"abc", 12) for k in range(5)]
    p_eq = ctx.plan(hb.HostPlan.from_yaml([datagen.rule_yaml([3], [])]))
    p_sim = ctx.plan(hb.HostPlan.from_yaml(PLANS["two_eq_two_sim"]))
    specs = [(p_eq, big)] + [(p_sim, s) for s in small]
    tables = [(ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))) for _, (L, R) in specs]
    try:
        jobs = ctx.submit_batch([(p, tl, tr) for (p, _), (tl, tr) in zip(specs, tables)])
        res = [_results(j) for j in jobs]
        assert len(res[0][0]) == 1500 * 1400 and res[0][2]["retries"] >= 1
        for (p, _), (tl, tr), r in zip(specs[1:], tables[1:], res[1:]):
            alone = _results(ctx.submit(p, tl, tr))
            H.assert_same_pairs(r[0], alone[0], "small job")
            H.assert_same_audit(r[1], alone[1], "small job")
    finally:
        for tl, tr in tables:
            tl.release()
            tr.release()
        p_eq.release()
        p_sim.release()


@pytest.mark.parametrize("shape", [(1, 256, 4), (4, 1024, 6)])
def test_filter_keeping_most_pairs(ctx, shape):
    """Join groups of ~650 x 600 rows where the bound filter keeps most pairs
    (two-letter strings, JW >= 0.5): work-groups fill their survivor buffers
    many times over and spill to the global list; the pairs still equal the
    oracle's, and a second batch of the same plan sizes its lists from the
    first (no rerun)."""
    L, R = _adversarial_tables(77, 1300, 1200, b"ab", 30, n_keys=2)
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.5)])])
    old = ctx.filter_shape
    ctx.filter_shape = shape
    dp = ctx.plan(plan)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    try:
        want, want_audit, _ = H.run_oracle(plan, L, R)
        first = _results(ctx.submit(dp, tl, tr))
        second = _results(ctx.submit(dp, tl, tr))
        assert first[2]["filtered"] > 0.5 * first[2]["candidates"]
        assert second[2]["retries"] == 0
        for got in (first, second):
            H.assert_same_pairs(got[0], want, "JW >= 0.5 on two letters")
            H.assert_same_audit(got[1], want_audit, "JW >= 0.5 on two letters")
    finally:
        ctx.filter_shape = old
        tl.release()
        tr.release()
        dp.release()


@pytest.mark.parametrize("rules", [
    [datagen.rule_yaml([3], [(1, 0.6)])],
    [datagen.rule_yaml([3, 4], [])],
    [datagen.rule_yaml([3], [(2, 0.5), (1, 0.55)])],
])
def test_many_tiny_join_groups(ctx, rules):
    """Join keys with ~2,000 values over 3,000 x 2,500 rows: thousands of
    groups of one or two rows a side (many left keys without a right
    partner), so the pairs of consecutive groups pack densely into the
    small-group path, in one job and across the jobs of a batch."""
    L, R = _adversarial_tables(91, 3000, 2500, b"abc", 16, n_keys=2000)
    plan = hb.HostPlan.from_yaml(rules)
    dp = ctx.plan(plan)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    parts = [(0, 700), (700, 701), (701, 3000)]
    slices = [tl.slice(b, e) for b, e in parts]
    try:
        want, want_audit, _ = H.run_oracle(plan, L, R)
        got = _results(ctx.submit(dp, tl, tr))
        H.assert_same_pairs(got[0], want, "tiny groups, one job")
        H.assert_same_audit(got[1], want_audit, "tiny groups, one job")
        jobs = ctx.submit_batch([(dp, sl, tr) for sl in slices] + [(dp, tl, tr)])
        res = [_results(j) for j in jobs]
        H.assert_same_pairs(res[-1][0], want, "tiny groups, in a batch")
        for (b, e), r in zip(parts, res):
            w, wa, _ = H.run_oracle(plan, [c[b:e] for c in L], R)
            H.assert_same_pairs(r[0], w, "tiny groups, slice [%d, %d)" % (b, e))
            H.assert_same_audit(r[1], wa, "tiny groups, slice [%d, %d)" % (b, e))
    finally:
        for sl in slices:
            sl.release()
        tl.release()
        tr.release()
        dp.release()


def test_multi_rule_dedupe_set_regrowth(ctx):
    """A two-rule job where all 100 x 1,400 pairs satisfy both rules: its
    dedupe set starts far smaller than the 140,000 distinct pairs, fills,
    and grows over reruns; every pair is reported once, as by the oracle."""
    rows_l = [(str(i).encode(), b"t", b"a", b"k", b"k") for i in range(100)]
    rows_r = [(str(i).encode(), b"t", b"a", b"k", b"k") for i in range(1400)]
    L, R = H.columns_from_rows(rows_l), H.columns_from_rows(rows_r)
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([3], []), datagen.rule_yaml([4], [])])
    dp = ctx.plan(plan)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    try:
        got, audit, stats = _results(ctx.submit(dp, tl, tr))
        want, _, _ = H.run_oracle(plan, L, R)
        assert len(want) == 140000 and stats["retries"] >= 1
        H.assert_same_pairs(got, want, "two all-matching rules")
    finally:
        tl.release()
        tr.release()
        dp.release()


def test_batch_results_block_equals_each_job(ctx):
    """hb_batch_results gives every job's pairs and audit entries in one
    block, with offsets, equal to each job's own hb_job_results."""
    specs = _jobs(ctx, 12, 17)
    plans = {n: ctx.plan(hb.HostPlan.from_yaml(PLANS[n])) for n in PLANS}
    tables = [(ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))) for _, L, R in specs]
    try:
        jobs = ctx.submit_batch([(plans[n], tl, tr) for (n, _, _), (tl, tr) in zip(specs, tables)])
        ids, pairs, po, audit, ao = ctx.batch_results(jobs[5])
        assert list(ids) == [j.id for j in jobs] and len(po) == len(ao) == len(jobs) + 1
        for k, j in enumerate(jobs):
            p, a, st = j.results()
            H.assert_same_pairs(hb.sort_pairs(pairs[po[k]:po[k + 1]]), hb.sort_pairs(p), "job %d" % k)
            H.assert_same_audit(audit[ao[k]:ao[k + 1]], a, "job %d" % k)
            assert st["matches"] == po[k + 1] - po[k]
        ctx.release_jobs(jobs)
        with pytest.raises(hb.HBError):
            jobs[0].stats()
    finally:
        for tl, tr in tables:
            tl.release()
            tr.release()
        for p in plans.values():
            p.release()


def test_release_before_completion(ctx):
    """Jobs released while still running (never waited for) give their
    memory back only once the device is done with it: the jobs submitted
    around them, and later ones reusing that memory, stay exact."""
    L, R, _, cfg = datagen.make("c1")
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    dp = ctx.plan(plan)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    try:
        want = _results(ctx.submit(dp, tl, tr))
        for _ in range(3):
            jobs = [ctx.submit(dp, tl, tr) for _ in range(6)]
            for j in jobs[1::2]:
                j.release()  # still running, most likely
            for j in jobs[0::2]:
                got = _results(j)
                H.assert_same_pairs(got[0], want[0], "neighbour of a released job")
                H.assert_same_audit(got[1], want[1], "neighbour of a released job")
    finally:
        tl.release()
        tr.release()
        dp.release()


@pytest.mark.parametrize("shape", [(2, 192, 16), (1, 256, 4), (4, 512, 8)])
def test_allowance_tables_and_computed_allowances_agree(shape):
    """The tile filter reads its length allowances from a table per first
    threshold, built once per context; a context holding 64 tables filters
    further thresholds with allowances computed in the kernel.  The same
    jobs in a context that has filled its tables (computed allowances, beside
    a job of a tabled threshold in the same batch) and in a fresh context
    (tables) keep the same number of candidates in the filter and give the
    oracle's pairs and audit entries, for each number of rows per lane."""
    L, R = _adversarial_tables(91, 700, 600, b"abcdefgh", 40, n_keys=1)
    thr_new = 0.77
    plan_new = hb.HostPlan.from_yaml([datagen.rule_yaml([], [(1, thr_new)])])
    plan_old = hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(2, 0.5)])])
    want_new = H.run_oracle(plan_new, L, R)
    want_old = H.run_oracle(plan_old, L, R)
    tiny = H.columns_from_rows([(b"1", b"ab", b"ab", b"0", b"0")])
    full, fresh = hb.Context(0), hb.Context(0)
    try:
        got = {}
        for name, c in (("full", full), ("fresh", fresh)):
            if name == "full":
                tl, tr = c.upload(hb.ColumnTable(tiny)), c.upload(hb.ColumnTable(tiny))
                for i in range(64):  # 0.5 first, as plan_old's threshold
                    p = c.plan(hb.HostPlan.from_yaml([datagen.rule_yaml([], [(1, 0.5 + 0.004 * i)])]))
                    _results(c.submit(p, tl, tr))
                    p.release()
                tl.release()
                tr.release()
            c.filter_shape = shape
            tl, tr = c.upload(hb.ColumnTable(L)), c.upload(hb.ColumnTable(R))
            pn, po = c.plan(plan_new), c.plan(plan_old)
            got[name] = [_results(j) for j in c.submit_batch([(pn, tl, tr), (po, tl, tr)])]
            for p in (pn, po):
                p.release()
            tl.release()
            tr.release()
        for name in got:
            for (pairs, audit, stats), want, what in zip(got[name], (want_new, want_old),
                                                         ("thr 0.77", "thr 0.5")):
                H.assert_same_pairs(pairs, want[0], "%s context, %s" % (name, what))
                H.assert_same_audit(audit, want[1], "%s context, %s" % (name, what))
        for k in range(2):
            assert got["full"][k][2]["filtered"] == got["fresh"][k][2]["filtered"]
        assert got["fresh"][0][2]["filtered"] < got["fresh"][0][2]["candidates"]
    finally:
        full.close()
        fresh.close()


@pytest.mark.parametrize("shape", [(1, 64, 4), (1, 256, 4), (2, 192, 16), (2, 512, 8),
                                   (4, 256, 8), (4, 1024, 6)])
def test_dense_flushes_with_partial_tiles(ctx, shape):
    """The filter's dense path (flushes of the work-group's survivor buffer
    between steps) for an equality-only rule and a JW >= 0.5 rule on two
    letters, with join groups of 700 x 600 rows, so that every group ends
    in a partly filled A tile (some waves idle) while the buffer fills and
    is flushed many times: pairs and audit entries equal the oracle's at
    each launch shape."""
    L, R = _adversarial_tables(131, 1400, 1200, b"ab", 24, n_keys=2)
    old = ctx.filter_shape
    ctx.filter_shape = shape
    plans = [hb.HostPlan.from_yaml([datagen.rule_yaml([3], [])]),
             hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.5)])]),
             hb.HostPlan.from_yaml([datagen.rule_yaml([3], []), datagen.rule_yaml([4], [(2, 0.5)])])]
    dps = [ctx.plan(p) for p in plans]
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    try:
        jobs = ctx.submit_batch([(dp, tl, tr) for dp in dps])
        for k, (plan, job) in enumerate(zip(plans, jobs)):
            got = _results(job)
            want, want_audit, _ = H.run_oracle(plan, L, R)
            H.assert_same_pairs(got[0], want, "plan %d, shape %s" % (k, shape))
            H.assert_same_audit(got[1], want_audit, "plan %d, shape %s" % (k, shape))
            assert got[2]["filtered"] > 0.3 * got[2]["candidates"]
    finally:
        ctx.filter_shape = old
        tl.release()
        tr.release()
        for dp in dps:
            dp.release()


def test_autotuner_picks_a_candidate_and_changes_no_result(ctx):
    """hb_autotune_filter_shape times the given shapes on a job sample and
    makes the fastest the context's: the chosen shape is one of the
    candidates, its time is positive, and jobs run afterwards still equal
    the oracle."""
    L, R = _adversarial_tables(141, 2500, 2200, b"abcdefgh", 40, n_keys=3)
    plan = hb.HostPlan.from_yaml(PLANS["two_eq_two_sim"])
    plan2 = hb.HostPlan.from_yaml(PLANS["eq_sim"])
    dp, dp2 = ctx.plan(plan), ctx.plan(plan2)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    old = ctx.filter_shape
    cands = [(1, 256, 4), (2, 192, 16), (2, 512, 8), (4, 256, 12)]
    try:
        best, ms = ctx.autotune([(dp, tl, tr), (dp2, tl, tr)], cands)
        assert best in cands and ms > 0
        assert ctx.filter_shape == best
        for p, d in ((plan, dp), (plan2, dp2)):
            got = _results(ctx.submit(d, tl, tr))
            want, want_audit, _ = H.run_oracle(p, L, R)
            H.assert_same_pairs(got[0], want, "after autotune %s" % (best,))
            H.assert_same_audit(got[1], want_audit, "after autotune %s" % (best,))
        best2, _ = ctx.autotune([(dp2, tl, tr)])  # the built-in candidate list
        assert ctx.filter_shape == best2
    finally:
        ctx.filter_shape = old
        tl.release()
        tr.release()
        dp.release()
        dp2.release()
