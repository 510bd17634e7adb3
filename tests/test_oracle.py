"""Pins the CPU oracle (test infrastructure) before it is trusted as the
checker of the HIP path.

The reference ships no tests, golden vectors or known-answer files for this
path, so the checks here come from what the mathematics fixes:
  * hand-derived known answers of lev_jaro_ratio / lev_jaro_winkler_ratio
    (core/gpu/device_func.cuh:54-132, 210-231);
  * two independently structured formulations of the greedy match (the
    reference's two loops, and a window + free-position bitmask) compared
    exhaustively on short strings over small alphabets and on long strings;
  * a closed form for strings of distinct bytes (no competition between equal
    bytes, so a position matches iff it lies in its window);
  * the greedy match decomposing per byte value;
  * the fp32/fp64 tail against numpy's IEEE arithmetic and, for the fused
    variant, exact rational arithmetic with one rounding;
  * the plan interpreter (core/gpu/global_func.cuh:118-167) against an
    OR-of-ANDs over every truth assignment of small plans;
  * the fixtures written by tests/golden/make_golden.py.
"""
import ctypes
import itertools
import json
import os
import random
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

CP = O.CHECK_POINT


def tail(m, S, T):
    """fp32 Jaro tail of device_func.cuh:113-128 in numpy IEEE arithmetic."""
    if m == 0:
        return np.float32(0.0)
    q1 = np.float32(m) / np.float32(S)
    q2 = np.float32(m) / np.float32(T)
    s12 = np.float32(q1 + q2)
    return np.float32((np.float64(s12) + 1.0) / 3.0)


# ---- known answers --------------------------------------------------------------

KATS = [
    # a, b, m, jaro (fp32 printed as double), p, jw
    (b"MARTHA", b"MARHTA", 6, 1.0, 3, 1.0),  # transpositions ignored
    (b"DIXON", b"DICKSONX", 4, 0.7666666507720947, 2, 0.8133333206176758),
    (b"DICKSONX", b"DIXON", 4, 0.7666666507720947, 2, 0.8133333206176758),
    (b"DWAYNE", b"DUANE", 4, 0.8222222328186035, 1, 0.8400000095367431),
    (b"CRATE", b"TRACE", 5, 1.0, 0, 1.0),  # wide window of the second loop
    (b"abcd", b"dcba", 2, 0.6666666865348816, 0, 0.6666666865348816),
    (b"abcdefghijXYZ", b"abcdefghijQRS", 10, 0.8461538553237915, 10, 1.0),  # clamped
    (b"abc", b"", 0, 0.0, 0, 0.0),
    (b"", b"", 0, 1.0, 0, 1.0),
]


@pytest.mark.parametrize("a,b,m,jaro,p,jw", KATS)
def test_known_answers(a, b, m, jaro, p, jw):
    assert O.jaro(a, b) == jaro
    assert O.jaro_windowed(a, b) == jaro
    assert O.jw(a, b) == jw
    if a and b:
        mm, S, T, pp = O.jaro_counts(a, b)
        assert (mm, pp) == (m, p)
        assert (S, T) == (min(len(a), len(b)), max(len(a), len(b)))


def test_thresholds_as_compared():
    # the plan holds fp32 thresholds; the kernel compares JW < (double)thr
    want = {0.89: 0.8899999856948853, 0.85: 0.8500000238418579, 0.81: 0.8100000023841858,
            0.79: 0.7900000214576721, 0.88: 0.8799999952316284, 0.80: 0.800000011920929,
            0.95: 0.949999988079071}
    for t, d in want.items():
        assert float(np.float32(t)) == d


# ---- greedy match --------------------------------------------------------------

def _all_strings(alphabet, max_len):
    for n in range(max_len + 1):
        for t in itertools.product(alphabet, repeat=n):
            yield bytes(t)


@pytest.mark.parametrize("alphabet,max_len", [(b"abc", 5), (b"abcd", 4), (b"ab", 7)])
def test_two_formulations_agree_exhaustive(alphabet, max_len):
    strs = list(_all_strings(alphabet, max_len))
    for a in strs:
        for b in strs:
            assert O.jaro(a, b) == O.jaro_windowed(a, b), (a, b)


def test_two_formulations_agree_long_strings():
    """Random strings up to 400 bytes, including inner strings of >= 257 bytes
    where the 256-entry index array's caps apply."""
    rng = random.Random(3)
    for it in range(1500):
        alpha = [b"ab", b"abcdef", bytes(range(1, 256))][it % 3]
        la, lb = rng.randint(0, 400), rng.randint(0, 400)
        if it % 5 == 0:
            la = lb = rng.randint(250, 300)
        a = bytes(rng.choice(alpha) for _ in range(la))
        b = bytes(rng.choice(alpha) for _ in range(lb))
        assert O.jaro(a, b) == O.jaro_windowed(a, b), (la, lb)


def _distinct_matches(s, t):
    """m for an inner string s and outer string t whose bytes are all distinct
    within each string: the byte at outer position i < n_out matches iff its
    position j in s lies in [max(0, i - h), min(i + h, S - 1)], capped at 255
    for i < h and 256 otherwise."""
    S, T = len(s), len(t)
    h = (S + 1) // 2
    n_out = min(T, S + h)
    pos = {c: j for j, c in enumerate(s)}
    m = 0
    for i in range(n_out):
        j = pos.get(t[i])
        if j is None:
            continue
        hi = min(i + h, S - 1, 255 if i < h else 256)
        if max(0, i - h) <= j <= hi:
            m += 1
    return m


def test_distinct_bytes_closed_form():
    rng = random.Random(5)
    pool = list(range(1, 256))
    for it in range(3000):
        la, lb = rng.randint(1, 120), rng.randint(1, 120)
        if it % 4 == 0:
            la = lb
        a = bytes(rng.sample(pool, la))
        # b shares a random subset of a's bytes at shuffled positions
        shared = rng.sample(list(a), rng.randint(0, min(la, lb)))
        rest = [c for c in pool if c not in a]
        b = shared + rng.sample(rest, lb - len(shared))
        rng.shuffle(b)
        b = bytes(b)
        s, t = (a, b) if la <= lb else (b, a)
        m, S, T, p = O.jaro_counts(a, b)
        assert m == _distinct_matches(s, t), (a, b)


def test_greedy_decomposes_per_byte_value():
    """Matches only pair equal bytes, so m(a, b) = sum over byte values c of
    m(a with every other byte replaced by x, b with every other byte replaced
    by y), x != y; the lengths, hence windows, are unchanged."""
    rng = random.Random(9)
    for _ in range(600):
        alpha = b"abcd"
        a = bytes(rng.choice(alpha) for _ in range(rng.randint(1, 60)))
        b = bytes(rng.choice(alpha) for _ in range(rng.randint(1, 60)))
        total = 0
        for c in set(a) & set(b):
            ma = bytes(x if x == c else ord("x") for x in a)
            mb = bytes(x if x == c else ord("y") for x in b)
            total += O.jaro_counts(ma, mb)[0]
        assert O.jaro_counts(a, b)[0] == total


def test_invariants():
    rng = random.Random(1)
    for _ in range(3000):
        alpha = rng.choice([b"ab", b"abc", b"abcdefghij", bytes(range(32, 127))])
        a = bytes(rng.choice(alpha) for _ in range(rng.randint(0, 90)))
        b = bytes(rng.choice(alpha) for _ in range(rng.randint(0, 90)))
        j, w = O.jaro(a, b), O.jw(a, b)
        assert 0.0 <= j <= 1.0 and 0.0 <= w <= 1.0
        assert w >= j
        if a:
            assert O.jaro(a, a) == 1.0 and O.jw(a, a) == 1.0
        if len(a) != len(b):
            assert O.jaro(a, b) == O.jaro(b, a)  # the shorter string is inner either way
        if not set(a) & set(b) and a and b:
            assert j == 0.0 and w == 0.0
        if a and b:
            m, S, T, p = O.jaro_counts(a, b)
            h = (S + 1) // 2
            assert m <= min(S, T, S + h)
            assert p == len(os.path.commonprefix([a, b]))
            assert j == tail(m, S, T)
            if p >= 10:
                assert w == 1.0


# ---- fp32 / fp64 tail ----------------------------------------------------------

def test_tail_counts_against_numpy_ieee():
    rng = random.Random(2)
    tup = [(m, S, T, p) for S in range(1, 40) for T in range(S, 60, 3)
           for m in range(0, S + 1, 2) for p in (0, 1, min(3, m), m)]
    for _ in range(5000):
        S = rng.randint(1, 300)
        T = rng.randint(S, 400)
        m = rng.randint(0, S)
        tup.append((m, S, T, rng.randint(0, m)))
    got = O.jw_from_counts_batch(tup)
    for (m, S, T, p), g in zip(tup, got):
        j = np.float64(tail(m, S, T))
        w = j + ((1.0 - j) * np.float64(p)) * 0.1  # numpy does not contract
        w = min(w, 1.0)
        assert g == w, (m, S, T, p)
        assert O.jw_from_counts(m, S, T, p) == g


def test_fused_variant_is_single_rounding():
    """hbo_jw_fused rounds j + ((1 - j) * p) * 0.1 once after the multiply:
    (1 - j) * p is exact for an fp32 j and p <= 256, so the exact rational
    value rounded once to double is the fused result."""
    rng = random.Random(4)
    checked = differs = 0
    for _ in range(4000):
        alpha = b"abcdef"
        a = bytes(rng.choice(alpha) for _ in range(rng.randint(1, 40)))
        b = a[: rng.randint(0, len(a))] + bytes(rng.choice(alpha) for _ in range(rng.randint(0, 30)))
        j = float(np.float32(O.jaro(a, b)))
        p = len(os.path.commonprefix([a, b]))
        exact = Fraction(j) + Fraction((1.0 - j) * p) * Fraction(0.1)
        want = min(float(exact), 1.0)
        assert O.jw_fused(a, b) == want
        checked += 1
        differs += O.jw_fused(a, b) != O.jw(a, b)
        assert abs(O.jw_fused(a, b) - O.jw(a, b)) <= 2.3e-16
    assert checked > 0


# ---- plan interpretation --------------------------------------------------------

def _truth_plan_tables(rules):
    """rules: list of lists of 'e'/'s'.  Predicate k of the plan reads column
    k; one left row and one right row per truth assignment, where a true
    predicate has equal cells ('e') or identical strings ('s', JW 1) and a false
    one has different cells ('e') or disjoint strings ('s', JW 0)."""
    preds = [t for r in rules for t in r]
    n = len(preds)
    assigns = list(itertools.product([False, True], repeat=n))
    idx, typ, thr = [], [], []
    k = 0
    for r in rules:
        for t in r:
            idx += [k, k]
            typ += [t, t]
            thr += [1.0, 1.0] if t == "e" else [0.5, 0.5]
            k += 1
        idx.append(CP)
        typ.append(" ")
        thr.append(1.0)
    lrows = [tuple(b"abcd" for _ in preds) for _ in assigns]
    rrows = [tuple((b"abcd" if v else (b"abce" if t == "e" else b"wxyz"))
                   for v, t in zip(a, preds)) for a in assigns]
    plan = O.Plan(len(rules), idx, "".join(typ).encode(), thr)
    return plan, H.columns_from_rows(lrows), H.columns_from_rows(rrows), assigns


@pytest.mark.parametrize("shape", [
    [["e"]], [["s"]], [["e", "s"]], [["s", "s", "e"]],
    [["e"], ["s"]], [["e", "e"], ["s"]], [["e", "s"], ["s", "e"]],
    [["s"], ["e", "s", "s"], ["e"]], [["e"], ["e"], ["s"], ["s", "e"]],
    [["e", "s", "e", "s", "e"]],
])
def test_plan_is_or_of_ands(shape):
    plan, L, R, assigns = _truth_plan_tables(shape)
    assert plan.check() == 0
    tl, tr = O.Table(L), O.Table(R)
    pairs, cnt = O.block(plan, tl, tr, n_threads=2)
    got = {(int(l), int(r)) for l, r in pairs}
    want = set()
    for ri, a in enumerate(assigns):
        # every left row holds the "true" cells; right row ri realises a
        rule_truth, k = [], 0
        for r in shape:
            rule_truth.append(all(a[k:k + len(r)]))
            k += len(r)
        if any(rule_truth):
            want |= {(li, ri) for li in range(len(assigns))}
    assert got == want


def test_plan_check_rejects_malformed():
    assert O.Plan(1, [0, 0, CP], b"ee ", [1, 1, 1]).check() == 0
    assert O.Plan(1, [0, 0], b"ee", [1, 1]).check() == -2       # no checkpoint
    assert O.Plan(1, [CP], b" ", [1]).check() == -3             # empty rule
    assert O.Plan(1, [0, 0, CP], b"xx ", [1, 1, 1]).check() == -4
    assert O.Plan(1, [0, 0, CP], b"es ", [1, 1, 1]).check() == -5
    assert O.Plan(1, [-1, 0, CP], b"ee ", [1, 1, 1]).check() == -6


def test_counters_follow_short_circuit_order():
    # rule: eq(col0) then sim(col1); only equal keys reach the JW
    L = H.columns_from_rows([(b"1", b"abc"), (b"2", b"abd")])
    R = H.columns_from_rows([(b"1", b"abc"), (b"1", b"xyz"), (b"3", b"abc")])
    plan = O.Plan(1, [0, 0, 1, 1, CP], b"eess ", [1, 1, 0.9, 0.9, 1])
    pairs, cnt = O.block(plan, O.Table(L), O.Table(R), n_threads=1)
    assert pairs.tolist() == [[0, 0]]
    assert cnt["pairs"] == 6 and cnt["eq_evals"] == 6 and cnt["jw_evals"] == 2
    assert cnt["matches"] == 1
    assert cnt["work"] == 6 + O.jw_work(b"abc", b"abc") + O.jw_work(b"abc", b"xyz")


def test_block_threads_and_sampling():
    rng = random.Random(8)
    rows = lambda n: [(str(i).encode(), bytes(rng.choice(b"abc") for _ in range(rng.randint(0, 9))),
                       str(rng.randrange(3)).encode()) for i in range(n)]
    L, R = H.columns_from_rows(rows(150)), H.columns_from_rows(rows(170))
    plan = O.Plan(2, [2, 2, 1, 1, CP, 1, 1, CP], b"eess ss ", [1, 1, .7, .7, 1, .9, .9, 1])
    tl, tr = O.Table(L), O.Table(R)
    p1, c1 = O.block(plan, tl, tr, n_threads=1)
    p8, c8 = O.block(plan, tl, tr, n_threads=8)
    assert np.array_equal(p1, p8) and c1 == c8
    ps, _ = O.block(plan, tl, tr, l_begin=2, l_step=7, n_threads=3)
    assert np.array_equal(ps, p1[(p1[:, 0] % 7) == 2])
    # per-pair entry point agrees with the blocked loop
    want = {(l, r) for l in range(0, 150, 13) for r in range(170)
            if O.lib().hbo_eval_pair(ctypes.byref(plan.view), ctypes.byref(tl.view), l,
                                     ctypes.byref(tr.view), r, None)}
    assert want == {tuple(x) for x in p1.tolist() if x[0] % 13 == 0}


def test_eval_pairs_agrees_with_block():
    """The batched per-pair evaluator gives exactly hbo_block's pair set on
    every pair of a small job."""
    rng = random.Random(9)
    rows = lambda n: [(str(i).encode(), bytes(rng.choice(b"abc") for _ in range(rng.randint(0, 9))),
                       str(rng.randrange(3)).encode()) for i in range(n)]
    L, R = H.columns_from_rows(rows(90)), H.columns_from_rows(rows(80))
    plan = O.Plan(2, [2, 2, 1, 1, CP, 1, 1, CP], b"eess ss ", [1, 1, .7, .7, 1, .9, .9, 1])
    tl, tr = O.Table(L), O.Table(R)
    pairs, _ = O.block(plan, tl, tr, n_threads=2)
    ll, rr = np.meshgrid(np.arange(90), np.arange(80), indexing="ij")
    every = np.stack([ll.ravel(), rr.ravel()], 1).astype(np.uint32)
    ok = O.eval_pairs(plan, tl, tr, every, n_threads=3)
    assert {tuple(x) for x in every[ok].tolist()} == {tuple(x) for x in pairs.tolist()}


# ---- equality and the audit band, pinned from first principles ----------------

@pytest.mark.parametrize("a,b,want", [
    (b"1", b"10", False),            # a prefix is not equal
    (b"10", b"1", False),
    (b"", b"", True),                # two empty cells
    (b"", b"0", False),
    (b"abc", b"abc", True),
    (b"abc", b"abd", False),
    (b"\x80\xff", b"\x80\xff", True),  # bytes above 127 compare as bytes
    (b"\x80\xff", b"\x80\xfe", False),
    (b"\xe9t\xe9", b"\xc3\xa9t\xc3\xa9", False),
])
def test_equal_direct_cases(a, b, want):
    """equalities_kernel (device_func.cuh:39-52): byte equality up to the
    first NUL, nothing more."""
    assert O.equal(a, b) is want
    assert O.equal(b, a) is want


def test_equal_stops_at_nul():
    """Cells are read up to their first NUL: bytes after it never count."""
    import ctypes as C
    a = C.create_string_buffer(b"ab\x00xy", 6)
    b = C.create_string_buffer(b"ab\x00zw", 6)
    assert O.lib().hbo_equal(a, b) == 1
    c = C.create_string_buffer(b"ab\x00", 4)
    d = C.create_string_buffer(b"abc", 4)
    assert O.lib().hbo_equal(c, d) == 0


def _plan_rules(plan):
    """The oracle plan's rules as lists of (slot, type, threshold)."""
    idx = list(plan.pred_index)
    typ = plan.pred_type.decode("latin-1")
    thr = list(plan.pred_threshold)
    rules, cur, i = [], [], 0
    while i < len(idx):
        if idx[i] == CP:
            rules.append(cur)
            cur = []
            i += 1
        else:
            cur.append((i, idx[i], idx[i + 1], typ[i], thr[i]))
            i += 2
    return rules


def _audit_brute_force(plan, L, R):
    """The audit list written out from its definition: for every pair and
    every rule on its own, predicates in plan order until the first false
    one (the reference's checkpoint skip, global_func.cuh:152-159); each
    similarity predicate evaluated there whose JW lies within 1e-6 of its
    threshold (compared as double, global_func.cuh:139-145) is an entry.
    Equality is Python byte equality; JW and its fused variant come from the
    oracle's functions pinned above; bit0 = the fused tail decides the
    threshold test the other way, bit1 = the shorter cell has >= 257 bytes."""
    out = []
    for l in range(len(L[0])):
        for r in range(len(R[0])):
            for rule in _plan_rules(plan):
                for slot, cl, cr, t, thr in rule:
                    a, b = bytes(L[cl][l]), bytes(R[cr][r])
                    if t == "e":
                        alive = a == b
                    else:
                        thr_d = float(np.float32(thr))
                        jw = O.jw(a, b)
                        if abs(jw - thr_d) < 1e-6:
                            flags = 1 if (jw < thr_d) != (O.jw_fused(a, b) < thr_d) else 0
                            if min(len(a), len(b)) >= 257:
                                flags |= 2
                            out.append((l, r, slot, flags, jw))
                        alive = not (jw < thr_d)
                    if not alive:
                        break
    out.sort()
    arr = np.zeros(len(out), dtype=O.AUDIT_DTYPE)
    for k, e in enumerate(out):
        arr[k] = e
    return arr


def _band_tables(seed, n, alphabet, max_len, long_len=0):
    rng = random.Random(seed)

    def s(m):
        return bytes(rng.choice(alphabet) for _ in range(rng.randint(0, m)))

    rows_l = [(str(i).encode(), s(max_len), s(max_len // 2 + 1), str(rng.randrange(2)).encode())
              for i in range(n)]
    rows_r = []
    for i in range(n):
        src = rows_l[rng.randrange(n)]
        t = bytearray(src[1])
        if t:
            t[rng.randrange(len(t))] = rng.choice(alphabet)
        rows_r.append((str(i).encode(), bytes(t) if rng.random() < 0.6 else s(max_len),
                       src[2] if rng.random() < 0.5 else s(max_len // 2 + 1),
                       str(rng.randrange(2)).encode()))
    if long_len:
        base = bytes(rng.choice(alphabet) for _ in range(long_len))
        rows_l[0] = (rows_l[0][0], base, rows_l[0][2], b"0")
        rows_r[0] = (rows_r[0][0], base[:-3] + b"xyz", rows_r[0][2], b"0")
        rows_r[1] = (rows_r[1][0], base[5:] + b"q" * 9, rows_r[1][2], b"0")
    return H.columns_from_rows(rows_l), H.columns_from_rows(rows_r)


@pytest.mark.parametrize("case", range(6))
def test_audit_list_brute_force(case):
    """hbo_audit against the audit definition written out in Python, on
    small tables whose thresholds are JW values of actual pairs (so the band
    holds entries), for plans of 1-3 rules mixing equalities and 1-2
    similarity predicates; one case holds 300-byte cells (flag bit1)."""
    alphabet = [b"ab", b"abc", b"abcd", bytes(range(1, 256)), b"ab", b"abc"][case]
    L, R = _band_tables(40 + case, 24, alphabet, 14, long_len=300 if case == 5 else 0)
    rng = random.Random(case)
    jws1 = sorted({O.jw(L[1][i], R[1][j]) for i in range(24) for j in range(24)} - {0.0, 1.0})
    jws2 = sorted({O.jw(L[2][i], R[2][j]) for i in range(24) for j in range(24)} - {0.0, 1.0})
    t1 = float(np.float32(jws1[rng.randrange(len(jws1))]))
    t2 = float(np.float32(jws2[rng.randrange(len(jws2))]))
    t3 = float(np.float32(jws1[len(jws1) // 3]))
    if case == 5:
        t1 = float(np.float32(O.jw(L[1][0], R[1][0])))
        t3 = float(np.float32(O.jw(L[1][0], R[1][1])))
    plans = [
        O.Plan(1, [1, 1, CP], b"ss ", [t1, t1, 1]),
        O.Plan(1, [3, 3, 1, 1, 2, 2, CP], b"eessss ", [1, 1, t1, t1, t2, t2, 1]),
        O.Plan(3, [3, 3, 2, 2, CP, 1, 1, CP, 2, 2, 1, 1, CP], b"eess ss ssss ",
               [1, 1, t2, t2, 1, t3, t3, 1, t2, t2, t1, t1, 1]),
    ]
    found = 0
    for plan in plans:
        assert plan.check() == 0
        want = _audit_brute_force(plan, L, R)
        got = O.audit(plan, O.Table(L), O.Table(R), n_threads=2)
        got = got[np.lexsort((got["slot"], got["r"], got["l"]))]
        H.assert_same_audit(got, want, "case %d" % case)
        found += len(want)
        if case == 5:
            assert np.any(want["flags"] & 2), "long cells give bit1 entries"
    assert found > 0


# ---- golden fixtures ------------------------------------------------------------

def _golden(name):
    with open(os.path.join(H.GOLDEN, name)) as f:
        return json.load(f)


def test_c1_golden():
    import hyperblocker_amd as hb
    from hyperblocker_amd import datagen
    L, R, dup, cfg = datagen.make("c1")
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    pairs, audit, cnt = H.run_oracle(plan, L, R)
    want = H.read_pairs(os.path.join(H.GOLDEN, "c1_pairs.txt"))
    assert np.array_equal(pairs, want)
    g = _golden("c1_digest.json")
    assert H.digest_pairs(pairs) == g["sha256"] and cnt == g["counters"]
    H.assert_same_audit(audit, H.audit_from_json(g["audit_entries"]), "c1")
    # every duplicated left row is found with its perturbed right copy
    assert {(int(l), int(r)) for l, r in pairs} == {(int(s), int(i)) for i, s in enumerate(dup)}


def test_c2_sample_golden():
    import hyperblocker_amd as hb
    from hyperblocker_amd import datagen
    L, R, dup, cfg = datagen.make("c2")
    plan = hb.HostPlan.from_yaml(cfg["rules"])
    pairs, audit, cnt = H.run_oracle(plan, L, R, l_step=97)
    want = H.read_pairs(os.path.join(H.GOLDEN, "c2_sample_pairs.txt"))
    assert np.array_equal(pairs, want)
    assert len(audit) == 0


@pytest.mark.parametrize("name", sorted(H.FIXTURES))
def test_csv_fixture_golden(name):
    Lh, Rh, plan = H.load_fixture(name)
    g = _golden(name + ".json")
    assert [Lh.n_rows, Rh.n_rows] == g["rows"]
    assert plan.pred_index == g["plan"]["pred_index"]
    assert plan.pred_type == g["plan"]["pred_type"]
    pairs, audit, cnt = H.run_oracle(plan, H.host_columns(Lh), H.host_columns(Rh))
    assert pairs.tolist() == g["pairs"] and H.digest_pairs(pairs) == g["sha256"]
    assert cnt == g["counters"]
    H.assert_same_audit(audit, H.audit_from_json(g["audit_entries"]), name)
    # the header line is data: its self-match "id,id" is emitted (row 0, row 0)
    assert [0, 0] in g["pairs"]
