"""Shared test helpers: build tables, run the HIP path and the CPU oracle on
the same inputs, compare results."""
import hashlib
import os

import numpy as np

import hyperblocker_amd as hb
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def columns_from_rows(rows):
    """rows: list of tuples of bytes -> list of numpy 'S' columns (itemsize
    > longest cell)."""
    n_cols = len(rows[0])
    cols = []
    for c in range(n_cols):
        vals = [r[c] for r in rows]
        w = max([len(v) for v in vals] + [0]) + 1
        cols.append(np.array(vals, dtype="S%d" % w))
    return cols


def host_columns(table):
    """HostTable (CSV read by the product reader) -> numpy 'S' columns."""
    return columns_from_rows([tuple(table.cell(r, c) for c in range(table.n_cols))
                              for r in range(table.n_rows)])


FIXTURES = {
    # name: (left csv, right csv, rule dir), all under tests/golden
    "dblp_acm": ("dblp_acm/tableA.csv", "dblp_acm/tableB.csv", "dblp_acm/rules"),
    "dblp_google_self": ("dblp_google/tableA.csv", "dblp_google/tableA.csv",
                         "dblp_google/rules"),
}


def load_fixture(name):
    """(HostTable L, HostTable R, HostPlan) of a CSV fixture."""
    lp, rp, rd = (os.path.join(GOLDEN, x) for x in FIXTURES[name])
    return hb.HostTable.from_csv(lp), hb.HostTable.from_csv(rp), hb.HostPlan.from_rule_dir(rd)


def read_pairs(path):
    a = np.loadtxt(path, dtype=np.uint32, ndmin=2)
    return a.reshape(-1, 2)


def audit_to_json(a):
    return [[int(e["l"]), int(e["r"]), int(e["slot"]), int(e["flags"]), float(e["jw"]).hex()]
            for e in sort_audit(a)]


def audit_from_json(rows):
    out = np.zeros(len(rows), dtype=O.AUDIT_DTYPE)
    for i, (l, r, slot, flags, jw) in enumerate(rows):
        out[i] = (l, r, slot, flags, float.fromhex(jw))
    return out


def oracle_plan(plan):
    return O.Plan(plan.n_rules, plan.pred_index, plan.pred_type.encode(), plan.pred_threshold)


def run_oracle(plan, L, R, l_step=1, audit=True):
    op = oracle_plan(plan)
    tl, tr = O.Table(L), O.Table(R)
    pairs, cnt = O.block(op, tl, tr, l_step=l_step)
    au = O.audit(op, tl, tr, l_step=l_step) if audit else None
    return pairs, au, cnt


def run_gpu(ctx, plan, L, R):
    tl = ctx.upload(hb.ColumnTable(L))
    tr = ctx.upload(hb.ColumnTable(R))
    dp = ctx.plan(plan)
    try:
        pairs, audit, stats = ctx.block(dp, tl, tr)
    finally:
        dp.release()
        tl.release()
        tr.release()
    return hb.sort_pairs(pairs), sort_audit(audit), stats


def sort_audit(a):
    if len(a) == 0:
        return a
    return a[np.lexsort((a["slot"], a["r"], a["l"]))]


def digest_pairs(sorted_pairs):
    arr = np.ascontiguousarray(np.asarray(sorted_pairs, dtype="<u4").reshape(-1, 2))
    return hashlib.sha256(arr.tobytes()).hexdigest()


def assert_same_pairs(got, want, what=""):
    got = hb.sort_pairs(got)
    want = hb.sort_pairs(want)
    if got.shape != want.shape or not np.array_equal(got, want):
        gs = set(map(tuple, got.tolist()))
        ws = set(map(tuple, want.tolist()))
        extra = sorted(gs - ws)[:10]
        missing = sorted(ws - gs)[:10]
        raise AssertionError("%s: %d pairs vs %d expected; extra %s missing %s"
                             % (what, len(gs), len(ws), extra, missing))


def assert_same_audit(got, want, what=""):
    got = sort_audit(got)
    want = sort_audit(np.asarray(want))
    assert len(got) == len(want), "%s: audit %d vs %d" % (what, len(got), len(want))
    for f in ("l", "r", "slot", "flags"):
        assert np.array_equal(got[f], want[f]), "%s: audit field %s differs" % (what, f)
    assert np.array_equal(got["jw"].view(np.uint64), want["jw"].view(np.uint64)), \
        "%s: audit JW bits differ" % what
