"""hb_block_noalloc against hb_block (needs an MI355X).

The no-allocation entry point runs a prepared job with all working memory
in the library's static scratch array and writes into the caller's device
buffers: its pairs and audit entries must equal hb_block's, on two calls in
a row (the second with a smaller input), for a one-rule plan and a
three-rule plan (dedupe sets in the scratch), and it must report outputs
beyond the caller's capacity and unprepared tables as errors.
"""
import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import capi, datagen
from tests import helpers as H
from tests.test_gpu_parity import PLANS, _adversarial_tables

pytestmark = pytest.mark.gpu


def _noalloc(ctx, dp, tl, tr, cap=400000, acap=100000):
    pb = hb.DeviceBytes(np.zeros(cap * 8, np.uint8))
    ab = hb.DeviceBytes(np.zeros(acap * capi.AUDIT_DTYPE.itemsize, np.uint8))
    try:
        n, na = ctx.block_noalloc(dp, tl, tr, pb, cap, ab, acap)
        pairs = pb.to_host(n * 8).view(np.uint32).reshape(-1, 2)
        audit = ab.to_host(na * capi.AUDIT_DTYPE.itemsize).view(capi.AUDIT_DTYPE)
        return hb.sort_pairs(pairs), H.sort_audit(audit)
    finally:
        pb.free()
        ab.free()


def test_noalloc_equals_block_on_two_calls():
    ctx = hb.Context(0)
    try:
        L, R, _, cfg = datagen.make("c1")
        plan = hb.HostPlan.from_yaml(cfg["rules"])
        dp = ctx.plan(plan)
        tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
        small = tl.slice(0, 700)
        try:
            with pytest.raises(hb.HBError) as e:  # not prepared yet
                ctx.block_noalloc(dp, tl, tr, None, 0, None, 0)
            assert e.value.status == capi.HB_ERR_INVALID_ARG
            for k, (a, b) in enumerate([(tl, tr), (small, tr)]):
                want, want_audit, _ = ctx.block(dp, a, b)  # prepares the tables
                got, got_audit = _noalloc(ctx, dp, a, b)
                H.assert_same_pairs(got, want, "call %d" % k)
                H.assert_same_audit(got_audit, H.sort_audit(want_audit), "call %d" % k)
                assert len(got) > 0
            # outputs beyond the caller's capacity: an error with the counts
            want, _, _ = ctx.block(dp, tl, tr)
            pb = hb.DeviceBytes(np.zeros(8 * 10, np.uint8))
            try:
                with pytest.raises(hb.HBError) as e:
                    ctx.block_noalloc(dp, tl, tr, pb, 10, None, 0)
                assert e.value.status == capi.HB_ERR_CAPACITY
                assert e.value.counts[0] == len(want)
            finally:
                pb.free()
        finally:
            small.release()
            tl.release()
            tr.release()
            dp.release()
    finally:
        ctx.close()


def test_noalloc_multi_rule_plan_matches_oracle():
    ctx = hb.Context(0)
    try:
        L, R = _adversarial_tables(61, 400, 300, b"abcd", 30, n_keys=3)
        plan = hb.HostPlan.from_yaml(PLANS["multi_rule"])
        dp = ctx.plan(plan)
        tl, tr = ctx.upload(hb.ColumnTable(L)), ctx.upload(hb.ColumnTable(R))
        try:
            ctx.block(dp, tl, tr)
            want, want_audit, _ = H.run_oracle(plan, L, R)
            for k in range(2):
                got, got_audit = _noalloc(ctx, dp, tl, tr, cap=200000, acap=50000)
                H.assert_same_pairs(got, want, "multi-rule call %d" % k)
                H.assert_same_audit(got_audit, want_audit, "multi-rule call %d" % k)
        finally:
            tl.release()
            tr.release()
            dp.release()
    finally:
        ctx.close()
