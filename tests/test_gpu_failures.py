"""Failure handling of the job paths on the GPU: a batch launch that fails
before any device work (forced from the host with the context's test hook
HB_INJECT_LAUNCH_FAILURES) reports its error through every job it held,
waits return instead of blocking, and the context and the serving loop keep
working for later jobs.  No kernel is made to fail."""
import os

import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import datagen
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _ctx_with_failures(n):
    os.environ["HB_INJECT_LAUNCH_FAILURES"] = str(n)
    try:
        return hb.Context(0)
    finally:
        del os.environ["HB_INJECT_LAUNCH_FAILURES"]


def test_failed_launch_reports_through_jobs_and_server_survives():
    L, R, _, cfg = datagen.make("c1", scale=(500, 400))
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    want, want_audit, _ = H.run_oracle(plan, L, R)
    ctx = _ctx_with_failures(2)
    try:
        dp = ctx.plan(plan)
        tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
        # plain submit: the error is returned, and the job reports it too
        with pytest.raises(hb.HBError) as e:
            ctx.submit(dp, tl, tr, job_id=901)
        assert e.value.status == hb.capi.HB_ERR_HIP
        job = hb.Job(ctx, 901)
        with pytest.raises(hb.HBError):
            job.wait()
        job.release()
        # served: the failed batch's job reports the error; the next one runs
        with hb.Server(ctx, max_batches_in_flight=2, max_jobs_per_batch=1) as srv:
            bad = srv.submit(dp, tl, tr)
            with pytest.raises(hb.HBError):
                bad.wait()
            bad.release()
            good = srv.submit(dp, tl, tr)
            pairs, audit, _ = good.results()
            good.release()
        H.assert_same_pairs(pairs, want, "served job after a failed launch")
        H.assert_same_audit(H.sort_audit(audit), want_audit, "served job after a failed launch")
        pairs, audit, _ = ctx.block(dp, tl, tr)
        H.assert_same_pairs(pairs, want, "plain job after failed launches")
        tl.release()
        tr.release()
        dp.release()
    finally:
        ctx.close()
