"""Captured batches (hb_batch_graph_*) on the HIP path (needs an MI355X).

A job set is prepared once and its launch sequence captured into a HIP
graph; every run replays it.  Each replay, with other GPU work and
synchronisations between the replays, must give every job the same pairs
and audit entries as the same jobs run eagerly.
"""
import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import datagen
from tests import helpers as H
from tests.test_gpu_batch import _jobs, _results
from tests.test_gpu_parity import PLANS, _adversarial_tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hb.Context(0)
    yield c
    c.close()


def _eager_between(ctx, plan, tables):
    """Other work on the GPU between replays: an eager batch, waited for."""
    for j in ctx.submit_batch([(plan, tl, tr) for tl, tr in tables]):
        _results(j)
    ctx.synchronize()


def test_graph_replays_equal_eager_batch(ctx):
    specs = _jobs(ctx, 30, 11)
    plans = {n: ctx.plan(hb.HostPlan.from_yaml(PLANS[n])) for n in PLANS}
    tables = [(ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))) for _, L, R in specs]
    jobs = [(plans[n], tl, tr) for (n, _, _), (tl, tr) in zip(specs, tables)]
    graph = None
    try:
        eager = [_results(j) for j in ctx.submit_batch(jobs)]
        graph = ctx.batch_graph(jobs)
        for rep in range(4):
            runs = graph.run()
            got = [_results(j) for j in runs]
            for k, ((name, L, R), e, g) in enumerate(zip(specs, eager, got)):
                what = "replay %d, job %d (%s, %d x %d)" % (rep, k, name, len(L[0]), len(R[0]))
                H.assert_same_pairs(g[0], e[0], what)
                H.assert_same_audit(g[1], e[1], what)
                assert g[2]["pairs"] == e[2]["pairs"] and g[2]["matches"] == len(e[0])
                assert g[2]["device_ms"] > 0
            _eager_between(ctx, plans[specs[0][0]], tables[:3])
    finally:
        if graph is not None:
            graph.release()
        for tl, tr in tables:
            tl.release()
            tr.release()
        for p in plans.values():
            p.release()


def test_graph_of_one_job(ctx):
    """A one-job graph (its pairs written in place by the exact stage),
    replayed without reading the runs between replays except the last: each
    run completes the previous one, whose results stay readable."""
    L, R, _, cfg = datagen.make("c1")
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    dp = ctx.plan(plan)
    tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
    graph = None
    try:
        want = _results(ctx.submit(dp, tl, tr))
        oracle, oracle_audit, _ = H.run_oracle(plan, L, R)
        H.assert_same_pairs(want[0], oracle, "c1 eager vs oracle")
        graph = ctx.batch_graph([(dp, tl, tr)])
        runs = [graph.run()[0] for _ in range(3)]
        for k, job in enumerate(runs):
            got = _results(job)
            H.assert_same_pairs(got[0], want[0], "c1 graph run %d" % k)
            H.assert_same_audit(got[1], want[1], "c1 graph run %d" % k)
    finally:
        if graph is not None:
            graph.release()
        tl.release()
        tr.release()
        dp.release()


def test_graph_run_that_overflows_is_captured_again(ctx):
    """A job emitting 2.1M pairs, with a plan whose hints are still small:
    the first run overflows the captured buffers and reruns eagerly with
    larger ones; the next run captures the graph again at those sizes and
    needs no rerun.  Every run's pairs are exact."""
    rows_l = [(str(i).encode(), b"x", b"y", b"k", b"k") for i in range(1500)]
    rows_r = [(str(i).encode(), b"x", b"y", b"k", b"k") for i in range(1400)]
    big = (H.columns_from_rows(rows_l), H.columns_from_rows(rows_r))
    small = _adversarial_tables(77, 100, 90, b"abc", 12)
    p_eq = ctx.plan(hb.HostPlan.from_yaml([datagen.rule_yaml([3], [])]))
    p_sim = ctx.plan(hb.HostPlan.from_yaml(PLANS["two_eq_two_sim"]))
    specs = [(p_eq, big), (p_sim, small)]
    tables = [(ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))) for _, (L, R) in specs]
    jobs = [(p, tl, tr) for (p, _), (tl, tr) in zip(specs, tables)]
    graph = None
    try:
        graph = ctx.batch_graph(jobs)
        first = [_results(j) for j in graph.run()]
        assert len(first[0][0]) == 1500 * 1400 and first[0][2]["retries"] >= 1
        alone = _results(ctx.submit(p_sim, *tables[1]))
        H.assert_same_pairs(first[1][0], alone[0], "small job, first run")
        H.assert_same_audit(first[1][1], alone[1], "small job, first run")
        for rep in range(2):
            again = [_results(j) for j in graph.run()]
            assert again[0][2]["retries"] == 0, "the recaptured graph should not overflow"
            H.assert_same_pairs(again[0][0], first[0][0], "big job, run %d" % (rep + 2))
            H.assert_same_pairs(again[1][0], alone[0], "small job, run %d" % (rep + 2))
            H.assert_same_audit(again[1][1], alone[1], "small job, run %d" % (rep + 2))
    finally:
        if graph is not None:
            graph.release()
        for tl, tr in tables:
            tl.release()
            tr.release()
        p_eq.release()
        p_sim.release()
