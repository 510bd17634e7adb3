"""Tables handed over in the reference's row-major layout, and the device
dictionary's exact fallback (needs an MI355X).

hb_table_upload_serialized (host rows) and hb_table_from_device (rows
already in device memory: blocking_kernel's d_tb_data with col_size /
col_offset, core/gpu/global_func.cuh:104-110) unpack the rows on the
device.  A cell is its bytes up to the first NUL or col_size bytes (the
reference's cells are zero padded, core/data_structures/table.h:112-126).

Equality codes come from a dictionary in device memory; rows it cannot
code exactly (a hash collision, a full table, a full arena) are coded by the
host.  Contexts created with a hash cut to 3 bits, 16 slots probed twice,
or no arena force each of those paths: the results stay the oracle's.
"""
import ctypes as C
import json
import os
import random

import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import capi, datagen
from tests import helpers as H
from tests.test_gpu_parity import PLANS, _adversarial_tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hb.Context(0)
    yield c
    c.close()


class _RawTable:
    """A row-major table with chosen col_size per column: cells longer than
    col_size - 1 fill their slot with no NUL, and `junk` bytes are written
    after the NUL of shorter cells (the reader must stop at the NUL)."""

    def __init__(self, columns, sizes, junk=b""):
        self.n_rows = len(columns[0])
        self.n_cols = len(columns)
        self.col_size = list(sizes)
        self.col_offset = [int(x) for x in np.concatenate([[0], np.cumsum(sizes)[:-1]])]
        self.aligned_tuple_size = (((sum(sizes) + 1) >> 6) << 6) + 64
        ats = self.aligned_tuple_size
        self.data = np.zeros(max(self.n_rows, 1) * ats, dtype=np.uint8)
        for c, (col, off, size) in enumerate(zip(columns, self.col_offset, sizes)):
            for r in range(self.n_rows):
                v = bytes(col[r])
                assert len(v) <= size
                cell = v + (b"\0" + junk if len(v) < size else b"")
                cell = cell[:size]
                self.data[r * ats + off: r * ats + off + len(cell)] = np.frombuffer(cell, np.uint8)
        self._size = (C.c_size_t * self.n_cols)(*self.col_size)
        self._offset = (C.c_size_t * self.n_cols)(*self.col_offset)
        self.view = capi.SerializedTable(self.n_rows, self.n_cols, ats, self._size, self._offset,
                                         self.data.ctypes.data)


def test_from_device_reference_layout(ctx):
    """dblp-acm in the reference layout, rows in device memory: the
    oracle's committed pairs and audit entries."""
    g = json.load(open(os.path.join(H.GOLDEN, "dblp_acm.json")))
    Lh, Rh, plan = H.load_fixture("dblp_acm")
    sl = hb.DeviceSerializedTable(hb.SerializedTable(H.host_columns(Lh)))
    sr = hb.DeviceSerializedTable(hb.SerializedTable(H.host_columns(Rh)))
    dp = ctx.plan(plan)
    try:
        tl, tr = ctx.from_device(sl), ctx.from_device(sr)
        got, audit, stats = ctx.block(dp, tl, tr)
        tl.release()
        tr.release()
        assert hb.sort_pairs(got).tolist() == g["pairs"]
        H.assert_same_audit(audit, H.audit_from_json(g["audit_entries"]), "from device")
    finally:
        dp.release()
        sl.free()
        sr.free()


@pytest.mark.parametrize("junk", [b"", b"zz\x01garbage"])
def test_unpack_cells_ending_at_nul_or_col_size(ctx, junk):
    """Cells filling their col_size with no NUL, junk after a NUL, empty
    cells and bytes above 127, through both row-major entry points: the
    pairs and audit entries of the oracle on the cells as defined."""
    rng = random.Random(7)
    alpha = b"ab\xe9c"
    rows = []
    for i in range(300):
        rows.append((str(i).encode(), bytes(rng.choice(alpha) for _ in range(rng.randint(0, 12))),
                     bytes(rng.choice(alpha) for _ in range(rng.randint(0, 7))),
                     str(rng.randrange(3)).encode()))
    cols = H.columns_from_rows(rows)
    sizes = [4, 12, 8, 1]  # ids up to 3 digits + NUL; title cells of 12 bytes fill their slot
    raw = _RawTable(cols, sizes, junk)
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.6), (2, 0.5)]),
                                  datagen.rule_yaml([], [(2, 0.9)])])
    want, want_audit, _ = H.run_oracle(plan, cols, cols)
    dp = ctx.plan(plan)
    dev = hb.DeviceSerializedTable(raw)
    try:
        for how in ("upload", "device"):
            t = ctx.upload_serialized(raw) if how == "upload" else ctx.from_device(dev)
            got, audit, _ = ctx.block(dp, t, t)
            t.release()
            H.assert_same_pairs(got, want, how)
            H.assert_same_audit(H.sort_audit(audit), want_audit, how)
    finally:
        dp.release()
        dev.free()


@pytest.mark.parametrize("n_extra", [20, 26])
def test_unpack_narrow_and_wide_tables(ctx, n_extra):
    """Tables of 24 columns (column parameters passed as kernel arguments)
    and of 30 (read from device memory), the extra columns of assorted
    sizes between the used ones: both row-major entry points give the
    oracle's pairs and audit entries."""
    rng = random.Random(11 + n_extra)
    alpha = b"abcd "
    rows = []
    for i in range(200):
        extra = tuple(bytes(rng.choice(alpha) for _ in range(rng.randint(0, 6)))
                      for _ in range(n_extra))
        rows.append((str(i).encode(), bytes(rng.choice(alpha) for _ in range(rng.randint(1, 10))),
                     bytes(rng.choice(alpha) for _ in range(rng.randint(0, 6))),
                     str(rng.randrange(4)).encode()) + extra)
    cols = H.columns_from_rows(rows)
    sizes = [4, 11, 7, 2] + [rng.randint(7, 12) for _ in range(n_extra)]
    raw = _RawTable(cols, sizes)
    assert raw.n_cols == 4 + n_extra
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.7), (2, 0.5)])])
    want, want_audit, _ = H.run_oracle(plan, cols, cols)
    dp = ctx.plan(plan)
    dev = hb.DeviceSerializedTable(raw)
    try:
        for how in ("upload", "device"):
            t = ctx.upload_serialized(raw) if how == "upload" else ctx.from_device(dev)
            got, audit, _ = ctx.block(dp, t, t)
            t.release()
            H.assert_same_pairs(got, want, how)
            H.assert_same_audit(H.sort_audit(audit), want_audit, how)
    finally:
        dp.release()
        dev.free()


@pytest.mark.parametrize("n_pick", [0, 1, 217])
def test_serialized_partition_rows(ctx, n_pick):
    """hb_table_upload_serialized_rows: rows listed out of order (a
    partition of the reference's serialized table) give the oracle's pairs
    and audit entries on those rows, in the listed order; a row index past
    the table is an error."""
    rng = np.random.default_rng(n_pick)
    alpha = b"abcde\xe9"
    rows = [(str(i).encode(), bytes(alpha[j] for j in rng.integers(0, 6, rng.integers(0, 15))),
             bytes(alpha[j] for j in rng.integers(0, 6, rng.integers(0, 9))),
             str(int(rng.integers(0, 4))).encode()) for i in range(400)]
    cols = H.columns_from_rows(rows)
    raw = _RawTable(cols, [4, 14, 9, 2], b"q\x02")
    pick_l = rng.permutation(400)[:n_pick]
    pick_r = rng.permutation(400)[:max(n_pick - 17, 0)]
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.6), (2, 0.5)]),
                                  datagen.rule_yaml([], [(1, 0.85)])])
    want, want_audit, _ = H.run_oracle(plan, [c[pick_l] for c in cols],
                                       [c[pick_r] for c in cols])
    dp = ctx.plan(plan)
    try:
        tl = ctx.upload_serialized(raw, rows=pick_l)
        tr = ctx.upload_serialized(raw, rows=pick_r)
        got, audit, _ = ctx.block(dp, tl, tr)
        tl.release()
        tr.release()
        H.assert_same_pairs(got, want, "partition rows")
        H.assert_same_audit(H.sort_audit(audit), want_audit, "partition rows")
        with pytest.raises(hb.HBError) as e:
            ctx.upload_serialized(raw, rows=[0, 400])
        assert e.value.status == capi.HB_ERR_INVALID_ARG
    finally:
        dp.release()


def test_serialized_table_rejects_bad_layout(ctx):
    raw = _RawTable(H.columns_from_rows([(b"1", b"ab")]), [2, 3])
    raw._size[1] = 200  # past aligned_tuple_size
    with pytest.raises(hb.HBError) as e:
        ctx.upload_serialized(raw)
    assert e.value.status == capi.HB_ERR_INVALID_ARG


def _ctx_with(env):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return hb.Context(0)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("env", [
    {"HB_DICT_HASH_BITS": "3"},                          # collisions everywhere
    {"HB_DICT_SLOTS": "4", "HB_DICT_MAX_PROBE": "2"},    # a full table
    {"HB_DICT_ARENA_MB": "0"},                           # no room for representatives
])
def test_dictionary_fallback_stays_exact(env):
    """Join keys of 2,000 values and composite keys, over two table pairs
    in one context (codes must agree across tables, whichever dictionary
    gave them): pairs and audit entries equal the oracle's."""
    c = _ctx_with(env)
    try:
        for seed in (91, 92):
            L, R = _adversarial_tables(seed, 1500, 1300, b"abc", 16, n_keys=2000)
            tl, tr = c.upload(hb.ColumnTable(L)), c.upload(hb.ColumnTable(R))
            for name in ("eq_only", "eq_sim", "two_eq_two_sim", "multi_rule"):
                plan = hb.HostPlan.from_yaml(PLANS[name])
                dp = c.plan(plan)
                got, audit, _ = c.block(dp, tl, tr)
                dp.release()
                want, want_audit, _ = H.run_oracle(plan, L, R)
                H.assert_same_pairs(got, want, "%s, %s, seed %d" % (env, name, seed))
                H.assert_same_audit(H.sort_audit(audit), want_audit, "%s, %s" % (env, name))
            tl.release()
            tr.release()
    finally:
        c.close()


@pytest.mark.parametrize("n_keys", [20, 45, 100])
def test_layouts_for_each_code_count(n_keys):
    """Join keys of n_keys values in a fresh context, so that the layouts
    see that many codes: up to 28 the counting layouts keep counts and
    offsets in LDS and alternate between two counter buffers, up to 64 they
    count with global atomics and scan once, beyond that radix sorts build
    them.  Three table pairs built one after another, each joined under
    two plans: pairs and audit entries equal the oracle's."""
    c = hb.Context(0)
    try:
        for seed in (11, 12, 13):
            L, R = _adversarial_tables(seed * 1000 + n_keys, 900, 800, b"abc", 16, n_keys=n_keys)
            tl, tr = c.upload(hb.ColumnTable(L)), c.upload(hb.ColumnTable(R))
            for name in ("eq_sim", "multi_rule"):
                plan = hb.HostPlan.from_yaml(PLANS[name])
                dp = c.plan(plan)
                got, audit, _ = c.block(dp, tl, tr)
                dp.release()
                want, want_audit, _ = H.run_oracle(plan, L, R)
                H.assert_same_pairs(got, want, "%d keys, %s, seed %d" % (n_keys, name, seed))
                H.assert_same_audit(H.sort_audit(audit), want_audit, "%d keys, %s" % (n_keys, name))
            tl.release()
            tr.release()
    finally:
        c.close()


def test_fused_column_build_equals_separate_kernels():
    """A similarity column's signatures, 128-class counters and bit-planes
    come from one fused kernel; a context created with
    HB_SEPARATE_COLUMN_BUILDS builds them with the three separate kernels.
    Jobs whose first thresholds go through the bound filter (0.89), the
    refine pass (0.6) and cells over 256 bytes give the same pairs, audit
    entries and filter / exact-stage counts in both."""
    L, R = _adversarial_tables(77, 900, 800, b"abcdefgh", 60, long_frac=0.05)
    plans = [hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.89), (2, 0.5)])]),
             hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.6)])]),
             hb.HostPlan.from_yaml(PLANS["multi_rule"])]
    got = {}
    for name, env in (("fused", {}), ("separate", {"HB_SEPARATE_COLUMN_BUILDS": "1"})):
        c = _ctx_with(env)
        try:
            tl, tr = c.upload(hb.ColumnTable(L)), c.upload(hb.ColumnTable(R))
            got[name] = []
            for plan in plans:
                dp = c.plan(plan)
                pairs, audit, st = c.block(dp, tl, tr)
                dp.release()
                got[name].append((hb.sort_pairs(pairs), H.sort_audit(audit), st))
            tl.release()
            tr.release()
        finally:
            c.close()
    for k, (a, b) in enumerate(zip(got["fused"], got["separate"])):
        H.assert_same_pairs(a[0], b[0], "plan %d" % k)
        H.assert_same_audit(a[1], b[1], "plan %d" % k)
        for key in ("candidates", "filtered", "jw_exact"):
            assert a[2][key] == b[2][key], (k, key, a[2][key], b[2][key])
    want, want_audit, _ = H.run_oracle(plans[1], L, R)
    H.assert_same_pairs(got["fused"][1][0], want, "refine plan vs oracle")


@pytest.mark.parametrize("plan_name", ["eq_sim", "two_eq_two_sim"])
@pytest.mark.parametrize("env", [{}, {"HB_DICT_HASH_BITS": "3"}])
def test_deferred_coding_reruns_when_codes_change(env, plan_name):
    """A key column (or list of key columns, coded as a composite) whose
    last codings added no value is coded without the host waiting; layouts
    built from it check every code against the count then known.  Fresh
    tables with the same key values run without a rerun; a table bringing
    new key values (and, with a 3-bit hash, rows the host must code) flags
    its layout, which is rebuilt, and the batch runs again: every result
    equals the oracle's."""
    c = _ctx_with(env)
    try:
        plan = hb.HostPlan.from_yaml(PLANS[plan_name])
        dp = c.plan(plan)
        same = [_adversarial_tables(300 + k, 700, 600, b"abcd", 20, n_keys=5) for k in range(4)]
        new = _adversarial_tables(399, 700, 600, b"abcd", 20, n_keys=60)
        retries = []
        for k, (L, R) in enumerate(same + [new, same[0]]):
            tl, tr = c.upload(hb.ColumnTable(L)), c.upload(hb.ColumnTable(R))
            got, audit, st = c.block(dp, tl, tr)
            tl.release()
            tr.release()
            want, want_audit, _ = H.run_oracle(plan, L, R)
            H.assert_same_pairs(got, want, "%s table pair %d" % (env, k))
            H.assert_same_audit(H.sort_audit(audit), want_audit, "%s table pair %d" % (env, k))
            retries.append(st["retries"])
        if not env:
            assert retries[4] >= 1, retries  # the new values were caught and the job rerun
            assert retries[2] == 0 and retries[3] == 0, retries
        dp.release()
    finally:
        c.close()


def test_deferred_coding_after_host_fallback():
    """A context whose device dictionary is full (64 slots): once some
    column's rows were coded by the host, a key column predicted stable
    (coded without waiting) that meets new values has rows still left to the
    host when its layout is built.  Those rows are flagged, the layout is
    rebuilt once their codes are final, and the results equal the oracle's,
    for a single key column and for a composite."""
    c = _ctx_with({"HB_DICT_SLOTS": "64", "HB_DICT_MAX_PROBE": "64"})
    try:
        few = [_adversarial_tables(500 + k, 500, 400, b"abcd", 20, n_keys=4) for k in range(2)]
        strings = _adversarial_tables(520, 500, 400, b"abcdefgh", 30)
        many = _adversarial_tables(530, 500, 400, b"abcd", 20, n_keys=150)
        steps = [(PLANS[p], t) for p in ("eq_sim", "two_eq_two_sim") for t in few]
        steps += [([datagen.rule_yaml([1], [])], strings)]            # host codes most of column 1
        steps += [(PLANS["eq_sim"], many), (PLANS["two_eq_two_sim"], many)]
        retries = []
        for k, (rules, (L, R)) in enumerate(steps):
            plan = hb.HostPlan.from_yaml(rules)
            dp = c.plan(plan)
            tl, tr = c.upload(hb.ColumnTable(L)), c.upload(hb.ColumnTable(R))
            got, audit, st = c.block(dp, tl, tr)
            tl.release()
            tr.release()
            dp.release()
            want, want_audit, _ = H.run_oracle(plan, L, R)
            H.assert_same_pairs(got, want, "step %d" % k)
            H.assert_same_audit(H.sort_audit(audit), want_audit, "step %d" % k)
            retries.append(st["retries"])
        assert retries[-2] >= 1 and retries[-1] >= 1, retries
    finally:
        c.close()
