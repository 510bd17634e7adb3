"""HIP path vs the CPU oracle on the same inputs (needs an MI355X).

Bar: identical pair sets and identical audit lists (bit-identical JW) for
integer/byte work; the JW tail bit-identical to the oracle's fp32/fp64
sequence.
"""
import json
import os
import random
import subprocess

import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import capi, datagen
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hb.Context(0)
    yield c
    c.close()


def _jw_device(ctx, tuples):
    import ctypes as C
    t = np.ascontiguousarray(np.asarray(tuples, dtype=np.int32).reshape(-1, 4))
    out = np.zeros(len(t), np.float64)
    outf = np.zeros(len(t), np.float64)
    capi.check(capi.lib().hb_jw_from_counts(
        ctx._h, t.ctypes.data_as(C.POINTER(C.c_int32)), len(t),
        out.ctypes.data_as(C.POINTER(C.c_double)), outf.ctypes.data_as(C.POINTER(C.c_double))))
    return out, outf


def test_jw_tail_bits_match_oracle(ctx):
    """Every (m, S, T, p) with S <= 48, T <= 72, plus random tuples up to 300:
    device JW == oracle JW bit for bit."""
    tup = [(m, S, T, p) for S in range(0, 49) for T in range(S, 73)
           for m in range(0, S + 1) for p in sorted({0, 1, 2, min(S, 4), min(S, 10), S})
           if p <= m]
    rng = random.Random(7)
    for _ in range(20000):
        S = rng.randint(1, 300)
        T = rng.randint(S, 400)
        m = rng.randint(0, S)
        tup.append((m, S, T, rng.randint(0, m)))
    tup = np.array(tup, dtype=np.int32)
    got, _ = _jw_device(ctx, tup)
    want = np.array([O.jw_from_counts(*map(int, t)) for t in tup])
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    assert len(bad) == 0, "first mismatches: %s" % [(tuple(tup[i]), got[i], want[i])
                                                    for i in bad[:5]]


def test_c1_parity(ctx):
    L, R, dup, cfg = datagen.make("c1")
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    got, audit, stats = H.run_gpu(ctx, plan, L, R)
    want, want_audit, cnt = H.run_oracle(plan, L, R)
    H.assert_same_pairs(got, want, "c1")
    H.assert_same_audit(audit, want_audit, "c1")
    assert stats["pairs"] == 2600 * 2300
    assert stats["matches"] == len(want) == 2200


def _adversarial_tables(seed, n_l, n_r, alphabet=b"abc", max_len=40, long_frac=0.0,
                        n_keys=3):
    rng = random.Random(seed)

    def s(maxl):
        n = rng.randint(0, maxl)
        return bytes(rng.choice(alphabet) for _ in range(n))

    def row(i):
        ln = 300 if rng.random() < long_frac else max_len
        return (str(i).encode(), s(ln), s(ln // 2 + 1),
                str(rng.randrange(n_keys)).encode(), str(rng.randrange(n_keys)).encode())

    Lr = [row(i) for i in range(n_l)]
    Rr = [row(i) for i in range(n_r)]
    # some near copies so similarity predicates fire at high thresholds
    for i in range(min(n_l, n_r) // 3):
        src = Lr[rng.randrange(n_l)]
        t = bytearray(src[1])
        if t and rng.random() < 0.5:
            t[rng.randrange(len(t))] = rng.choice(alphabet)
        Rr[i] = (Rr[i][0], bytes(t), src[2], src[3], src[4])
    return H.columns_from_rows(Lr), H.columns_from_rows(Rr)


PLANS = {
    "eq_sim": [datagen.rule_yaml([3], [(1, 0.7)])],
    "sim_only": [datagen.rule_yaml([], [(1, 0.8)])],
    "eq_only": [datagen.rule_yaml([3, 4], [])],
    "two_eq_two_sim": [datagen.rule_yaml([3, 4], [(1, 0.6), (2, 0.55)])],
    "multi_rule": [datagen.rule_yaml([3], [(1, 0.85)]),
                   datagen.rule_yaml([4], [(2, 0.7), (1, 0.5)]),
                   datagen.rule_yaml([], [(1, 0.93)])],
}


@pytest.mark.parametrize("name", sorted(PLANS))
@pytest.mark.parametrize("seed,alphabet,max_len,long_frac", [
    (1, b"abc", 12, 0.0),
    (2, b"abcdefgh", 40, 0.0),
    (3, bytes(range(1, 256)), 70, 0.0),
    (4, b"ab", 300, 0.2),
])
def test_adversarial_parity(ctx, name, seed, alphabet, max_len, long_frac):
    L, R = _adversarial_tables(seed, 160, 150, alphabet, max_len, long_frac)
    plan = hb.HostPlan.from_yaml(PLANS[name])
    got, audit, stats = H.run_gpu(ctx, plan, L, R)
    want, want_audit, _ = H.run_oracle(plan, L, R)
    H.assert_same_pairs(got, want, name)
    H.assert_same_audit(audit, want_audit, name)


FILTER_SHAPES = [(1, 256, 4), (2, 256, 16), (2, 512, 12), (4, 256, 8), (4, 1024, 6), (2, 64, 64),
                 (4, 192, 32)]


@pytest.mark.parametrize("shape", FILTER_SHAPES)
def test_filter_shapes_give_identical_results(ctx, shape):
    """Every launch shape of the bound filter (rows per lane, B tile,
    groups per CU) returns the oracle's pairs; join groups here are large
    enough for the tiled kernel, with odd-length runs of equal B length."""
    saved = ctx.filter_shape
    ctx.filter_shape = shape
    try:
        assert ctx.filter_shape == shape
        L, R = _adversarial_tables(21, 900, 700, b"abcdefgh", 40)
        for name in ("eq_sim", "sim_only", "two_eq_two_sim", "multi_rule"):
            plan = hb.HostPlan.from_yaml(PLANS[name])
            got, audit, stats = H.run_gpu(ctx, plan, L, R)
            want, want_audit, _ = H.run_oracle(plan, L, R)
            H.assert_same_pairs(got, want, "%s %s" % (name, shape))
            H.assert_same_audit(audit, want_audit, "%s %s" % (name, shape))
    finally:
        ctx.filter_shape = saved


def test_filter_shape_validation(ctx):
    for bad in [(3, 256, 8), (2, 100, 8), (2, 32, 16), (2, 256, 0), (2, 256, 65), (1, 131072, 4)]:
        with pytest.raises(capi.HBError) as e:
            ctx.filter_shape = bad
        assert e.value.status == capi.HB_ERR_INVALID_ARG


def test_audit_band_entries(ctx):
    """Thresholds set to JW values of actual pairs put those pairs in the
    audit band; the device lists exactly the oracle's entries."""
    L, R = _adversarial_tables(11, 60, 60, b"abcd", 16)
    vals = sorted({O.jw(L[1][i], R[1][j]) for i in range(0, 60, 7) for j in range(0, 60, 5)})
    thr = float(np.float32(vals[len(vals) // 2]))
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([], [(1, thr)])])
    got, audit, _ = H.run_gpu(ctx, plan, L, R)
    want, want_audit, _ = H.run_oracle(plan, L, R)
    H.assert_same_pairs(got, want, "audit")
    H.assert_same_audit(audit, want_audit, "audit")
    assert len(want_audit) > 0


def test_empty_and_single_row_tables(ctx):
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([3], [(1, 0.5)])])
    L, R = _adversarial_tables(5, 1, 40, b"ab", 5)
    got, audit, stats = H.run_gpu(ctx, plan, L, R)
    want, _, _ = H.run_oracle(plan, L, R)
    H.assert_same_pairs(got, want, "single")
    E = [np.zeros(0, dtype=c.dtype) for c in L]
    got, audit, stats = H.run_gpu(ctx, plan, E, R)
    assert len(got) == 0 and stats["pairs"] == 0


def test_capacity_regrowth(ctx):
    """An equality-only rule over few keys emits far more pairs than the
    initial buffers hold; the job regrows and returns all of them."""
    rows_l = [(str(i).encode(), b"x", b"y", b"k", b"k") for i in range(1500)]
    rows_r = [(str(i).encode(), b"x", b"y", b"k", b"k") for i in range(1400)]
    L, R = H.columns_from_rows(rows_l), H.columns_from_rows(rows_r)
    plan = hb.HostPlan.from_yaml([datagen.rule_yaml([3], [])])
    got, audit, stats = H.run_gpu(ctx, plan, L, R)
    assert len(got) == 1500 * 1400
    assert stats["retries"] >= 1
    ll, rr = np.meshgrid(np.arange(1500), np.arange(1400), indexing="ij")
    H.assert_same_pairs(got, np.stack([ll.ravel(), rr.ravel()], 1), "regrowth")


def test_cli_partitions_match_oracle(tmp_path):
    """run_hyperblocker_exec with -n_partitions 4: the output lines equal the
    oracle run on every partition of the host partitioner."""
    L, R, dup, cfg = datagen.make("c1", scale=(900, 800))
    d = datagen.write_dataset(str(tmp_path / "ds"), L, R, cfg["rules"])
    out = str(tmp_path / "out.csv")
    exe = os.path.join(H.ROOT, "hyperblocker_amd", "bin", "run_hyperblocker_exec")
    subprocess.run([exe, "-data_l", d + "/tableA.csv", "-data_r", d + "/tableB.csv",
                    "-rule_dir", d + "/rules", "-o", out, "-n_partitions", "4",
                    "-metrics", str(tmp_path / "m.json")], check=True,
                   stdout=subprocess.DEVNULL, timeout=120)
    got = sorted(open(out).read().splitlines())
    hl = hb.HostTable.from_csv(d + "/tableA.csv")
    hr = hb.HostTable.from_csv(d + "/tableB.csv")
    plan = hb.HostPlan.from_rule_dir(d + "/rules")
    want = []
    for lrows, rrows in hb.partition(hl, hr, plan, 4):
        Lp = [c[lrows] for c in L]
        Rp = [c[rrows] for c in R]
        pairs, _, _ = H.run_oracle(plan, Lp, Rp, audit=False)
        for l, r in pairs:
            want.append("%s,%s" % (L[0][lrows[l]].decode(), R[0][rrows[r]].decode()))
    assert got == sorted(want)
    m = json.load(open(tmp_path / "m.json"))
    assert m["matches"] == len(want)


def test_c2_full_size_digest(ctx):
    """Config 2 at full size (1e10 pairs): the sorted pair set hashes to the
    oracle's committed digest."""
    g = json.load(open(os.path.join(H.GOLDEN, "c2_digest.json")))
    L, R, dup, cfg = datagen.make("c2")
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    got, audit, stats = H.run_gpu(ctx, plan, L, R)
    assert len(got) == g["count"]
    assert H.digest_pairs(got) == g["sha256"]
    assert len(audit) == g["audit"]
    sample = H.read_pairs(os.path.join(H.GOLDEN, "c2_sample_pairs.txt"))
    assert np.array_equal(got[got[:, 0] % 97 == 0], sample)


@pytest.mark.parametrize("name", sorted(H.FIXTURES))
def test_csv_fixture_parity(ctx, name):
    """The reference's own data sets (dblp-acm; DBLP-Google tableA against
    itself), read by the product CSV reader and rule loader: pairs and audit
    list equal the oracle's committed results.  dblp-acm holds author cells of
    more than 256 bytes and DBLP-Google non-ASCII bytes."""
    g = json.load(open(os.path.join(H.GOLDEN, name + ".json")))
    Lh, Rh, plan = H.load_fixture(name)
    tl, tr = ctx.upload(Lh), ctx.upload(Rh)
    dp = ctx.plan(plan)
    try:
        got, audit, stats = ctx.block(dp, tl, tr)
    finally:
        dp.release()
        tl.release()
        tr.release()
    got = hb.sort_pairs(got)
    assert got.tolist() == g["pairs"]
    H.assert_same_audit(audit, H.audit_from_json(g["audit_entries"]), name)
    assert stats["pairs"] == g["counters"]["pairs"]


def test_cli_on_reference_dataset(tmp_path):
    """run_hyperblocker_exec on data/dblp-acm with its rule: the output lines
    are the oracle's pairs written as "<left id>,<right id>"."""
    d = os.path.join(H.GOLDEN, "dblp_acm")
    out = str(tmp_path / "out.csv")
    exe = os.path.join(H.ROOT, "hyperblocker_amd", "bin", "run_hyperblocker_exec")
    subprocess.run([exe, "-data_l", d + "/tableA.csv", "-data_r", d + "/tableB.csv",
                    "-rule_dir", d + "/rules", "-o", out, "-n_partitions", "1"], check=True,
                   stdout=subprocess.DEVNULL, timeout=120)
    g = json.load(open(os.path.join(H.GOLDEN, "dblp_acm.json")))
    Lh, Rh, _ = H.load_fixture("dblp_acm")
    want = sorted(Lh.cell(l, 0) + b"," + Rh.cell(r, 0) for l, r in g["pairs"])
    assert sorted(open(out, "rb").read().splitlines()) == want


def test_reference_layout_handoff(ctx):
    """The reference's row-major SerializedTable (table.h:75-128: col_size =
    longest cell + 1, aligned_tuple_size = (((sum + 1) >> 6) << 6) + 64,
    zero padding) of the dblp-acm tables, passed through
    hb_table_upload_serialized: pairs and audit entries equal the oracle's
    committed results for that data set."""
    g = json.load(open(os.path.join(H.GOLDEN, "dblp_acm.json")))
    Lh, Rh, plan = H.load_fixture("dblp_acm")
    sl, sr = hb.SerializedTable(H.host_columns(Lh)), hb.SerializedTable(H.host_columns(Rh))
    assert sl.col_offset[0] == 0 and sl.aligned_tuple_size % 64 == 0
    tl, tr = ctx.upload_serialized(sl), ctx.upload_serialized(sr)
    dp = ctx.plan(plan)
    try:
        got, audit, stats = ctx.block(dp, tl, tr)
    finally:
        dp.release()
        tl.release()
        tr.release()
    assert hb.sort_pairs(got).tolist() == g["pairs"]
    H.assert_same_audit(audit, H.audit_from_json(g["audit_entries"]), "serialized dblp_acm")
    assert stats["pairs"] == g["counters"]["pairs"]
