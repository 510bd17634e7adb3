"""Host-side logic of the drop-in front end, on the CPU: rule files -> plan
(core/common/yaml_config.h:19-57, core/components/execution_plan_generator.h:
64-155), CSV reading (core/components/data_mngr.h:46-79), partitioning
(data_mngr.h:81-222), the output writer (core/components/host_reducer.h:
80-90), the bound filter's coefficients, the gflags-compatible executable
(examples/run_hyperblocker.cu:22-59) up to the point where it needs a GPU,
and the C ABI library's exports.

The reference's CSV, YAML and flag libraries (rapidcsv, yaml-cpp, gflags) are
not in the image and ship no tests, so parity at that boundary is unpinned;
these tests pin this build's own behaviour, and the well-formed files
against Python's csv module.
"""
import csv
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import hyperblocker_amd as hb
from hyperblocker_amd import capi, datagen
from oracle import oracle as O
from tests import helpers as H

CP = capi.CHECK_POINT
HEADER = os.path.join(H.ROOT, "include", "hyperblocker_amd.h")


# ---- C ABI -------------------------------------------------------------------------

def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = _declared_functions()
    assert len(names) >= 30
    lib = C.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(capi.EXPORTS)


def test_no_device_is_an_error_not_a_fallback():
    """Without a HIP device the context cannot be created; nothing computes on
    the CPU instead."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    st = capi.lib().hb_ctx_create(0, C.byref(h))
    assert st == capi.HB_ERR_NO_DEVICE
    assert b"no HIP device" in capi.lib().hb_last_error()


def test_version_string():
    assert capi.lib().hb_version().startswith(b"hyperblocker_amd")


# ---- rules and plans ---------------------------------------------------------------

def test_reference_rule_files():
    p = hb.HostPlan.from_rule_dir(os.path.join(H.GOLDEN, "dblp_acm", "rules"))
    assert p.n_rules == 1
    assert p.pred_index == [4, 4, 1, 1, 2, 2, CP]
    assert p.pred_type == "eessss "
    assert p.pred_threshold == [1.0, 1.0, np.float32(0.81), np.float32(0.81),
                                np.float32(0.79), np.float32(0.79), 1.0]
    p = hb.HostPlan.from_rule_dir(os.path.join(H.GOLDEN, "dblp_google", "rules"))
    # Equalities [0, 1] holds only the relation ids: a rule of two JW predicates
    assert p.pred_index == [1, 1, 2, 2, CP] and p.pred_type == "ssss "
    assert p.pred_threshold == [np.float32(0.88)] * 2 + [np.float32(0.85)] * 2 + [1.0]


RULE = """Preconditions:
  Relations:
    - [0, 0, 1, 2, 3, 4]
    - [1, 0, 1, 2, 3, 4]
  Equalities:
    - {eq}
  Sim:
    - {sim}
  Threshold:
    - {thr}
Conseq:
  Option:
    Equality
  Equality:
    [0, 1, 0, 0]
"""


def _rule(eq, sim, thr):
    return RULE.format(eq=eq, sim=sim, thr=thr)


def test_rules_sorted_by_equality_count_and_flattened():
    docs = [_rule("[0, 1]", "[0, 1, 1, 1]", "[0, 1, 0.9]"),
            _rule("[0, 1, 4, 4, 3, 3]", "[0, 1, 2, 2]", "[0, 1, 0.7]"),
            _rule("[0, 1, 4, 4]", "[0, 1, 1, 1, 2, 2]", "[0, 1, 0.8, 0.6]")]
    p = hb.HostPlan.from_yaml(docs)
    assert p.n_rules == 3
    assert p.pred_index == [4, 4, 3, 3, 2, 2, CP, 4, 4, 1, 1, 2, 2, CP, 1, 1, CP]
    assert p.pred_type == "eeeess eessss ss "
    f = lambda x: float(np.float32(x))
    assert p.pred_threshold == [1, 1, 1, 1, f(.7), f(.7), 1, 1, 1, f(.8), f(.8), f(.6), f(.6), 1,
                                f(.9), f(.9), 1]


def test_threshold_keyed_by_left_column_first_insert_wins():
    f = lambda x: float(np.float32(x))
    # thresholds map left column -> value; the second insert for column 1 is
    # ignored, so both pairs on left column 1 get 0.9
    p = hb.HostPlan.from_yaml([_rule("[0, 1]", "[0, 1, 1, 1, 1, 1]", "[0, 1, 0.9, 0.5]")])
    assert p.pred_index == [1, 1, 1, 1, CP]
    assert p.pred_threshold == [f(.9)] * 4 + [1.0]
    # each slot looks its column up in that map, the right slot included; the
    # kernel compares against the left slot's value (global_func.cuh:139-145)
    p = hb.HostPlan.from_yaml([_rule("[0, 1]", "[0, 1, 1, 1, 2, 1]", "[0, 1, 0.9, 0.5]")])
    assert p.pred_index == [1, 1, 2, 1, CP]
    assert p.pred_threshold == [f(.9), f(.9), f(.5), f(.9), 1.0]
    # a right column with no entry (execution_plan_generator.h:130-131 would
    # read end()) is an error
    with pytest.raises(capi.HBError):
        hb.HostPlan.from_yaml([_rule("[0, 1]", "[0, 1, 1, 2]", "[0, 1, 0.9]")])


def test_yaml_flow_and_block_forms():
    block = """Preconditions:
  Relations:
    - [0, 0, 1]
    - [1, 0, 1]
  Equalities:
    -
      - 0
      - 1
      - 3
      - 3
  Sim:
    - [0, 1, 1, 1]
  Threshold:
    - ['0', "1", 0.85]
Conseq: {Option: Equality, Equality: [0, 1, 0, 0]}
"""
    p = hb.HostPlan.from_yaml([block])
    assert p.pred_index == [3, 3, 1, 1, CP] and p.pred_type == "eess "


@pytest.mark.parametrize("doc,status", [
    # Preconditions must have exactly 4 keys (yaml_config.h:30)
    ("Preconditions:\n  Relations:\n    - [0]\n    - [1]\n  Equalities:\n    - [0, 1, 4, 4]\n"
     "  Sim:\n    - [0, 1]\nConseq:\n  Option: Equality\n", capi.HB_ERR_PLAN),
    # Relations is read before that check; without it the document is malformed
    ("Preconditions:\n  Equalities:\n    - [0, 1, 4, 4]\n  Sim:\n    - [0, 1]\n"
     "  Threshold:\n    - [0, 1]\nConseq:\n  Option: Equality\n", capi.HB_ERR_PARSE),
    (_rule("[0, 1]", "[0, 1]", "[0, 1]"), capi.HB_ERR_PLAN),               # no predicate
    (_rule("[0, 1, 4]", "[0, 1]", "[0, 1]"), capi.HB_ERR_PLAN),            # odd column list
    (_rule("[0, 1]", "[0, 1, 1, 1]", "[0, 1, abc]"), capi.HB_ERR_PARSE),   # bad threshold
    (_rule("[0, 1]", "[0, 1, 1, 1]", "[0, 1, 0.9, 0.8]"), capi.HB_ERR_PLAN),  # extra threshold
    (_rule("[0, 1, -1, 2]", "[0, 1]", "[0, 1]"), capi.HB_ERR_PLAN),        # negative column
    ("Preconditions: [unclosed\n", capi.HB_ERR_PARSE),
])
def test_malformed_rules_are_rejected(doc, status):
    with pytest.raises(capi.HBError) as e:
        hb.HostPlan.from_yaml([doc])
    assert e.value.status == status


def test_rule_dir_errors(tmp_path):
    with pytest.raises(capi.HBError) as e:
        hb.HostPlan.from_rule_dir(str(tmp_path / "nope"))
    assert e.value.status == capi.HB_ERR_IO
    with pytest.raises(capi.HBError) as e:
        hb.HostPlan.from_rule_dir(str(tmp_path))
    assert e.value.status == capi.HB_ERR_PLAN


# ---- CSV ---------------------------------------------------------------------------

def _read_all(t):
    return [[t.cell(r, c) for c in range(t.n_cols)] for r in range(t.n_rows)]


def test_reference_csvs_match_python_csv():
    """Quoted CRLF files (dblp-acm) and an unquoted LF file with non-ASCII
    bytes (DBLP-Google): the header line is a data row."""
    for rel in ["dblp_acm/tableA.csv", "dblp_acm/tableB.csv", "dblp_google/tableA.csv"]:
        path = os.path.join(H.GOLDEN, rel)
        t = hb.HostTable.from_csv(path)
        with open(path, newline="", encoding="latin-1") as f:
            want = [[c.encode("latin-1") for c in row] for row in csv.reader(f)]
        assert _read_all(t) == want, rel


@pytest.mark.parametrize("data,sep,want", [
    (b"a,b\nc,d\n", ",", [[b"a", b"b"], [b"c", b"d"]]),
    (b"a,b\r\nc,d", ",", [[b"a", b"b"], [b"c", b"d"]]),                 # CRLF, no final newline
    (b"\xef\xbb\xbfa,b\n", ",", [[b"a", b"b"]]),                         # UTF-8 BOM skipped
    (b'"x,y",z\n"say ""hi""",w\n', ",", [[b"x,y", b"z"], [b'say "hi"', b"w"]]),
    (b"a,,c\n,,\n", ",", [[b"a", b"", b"c"], [b"", b"", b""]]),          # empty cells
    (b"a|b\nc|d\n", "|x", [[b"a", b"b"], [b"c", b"d"]]),                # first char of sep
    (b"a;b,c\n", ";", [[b"a", b"b,c"]]),
    (b"caf\xc3\xa9,\xff\n", ",", [[b"caf\xc3\xa9", b"\xff"]]),           # bytes kept as-is
])
def test_csv_edge_cases(tmp_path, data, sep, want):
    p = tmp_path / "t.csv"
    p.write_bytes(data)
    assert _read_all(hb.HostTable.from_csv(str(p), sep=sep)) == want


def _csv_line_model(line, sep):
    """rapidcsv's cell rules for one line (its LF removed): a quote toggles
    the quoted state only at the start of a cell or inside a cell that began
    with one, a separator outside quotes ends the cell, CR is dropped, and a
    cell "..." is unquoted with "" -> "."""
    def unquote(c):
        if len(c) >= 2 and c[:1] == b'"' and c[-1:] == b'"':
            out, k = bytearray(), 1
            while k + 1 < len(c):
                out.append(c[k])
                if c[k] == 0x22 and k + 2 < len(c) and c[k + 1] == 0x22:
                    k += 1
                k += 1
            return bytes(out)
        return bytes(c)
    cells, cell, quoted = [], bytearray(), False
    for ch in line:
        if ch == 0x22:
            if not cell or cell[0] == 0x22:
                quoted = not quoted
            cell.append(ch)
        elif ch == sep and not quoted:
            cells.append(unquote(cell))
            cell = bytearray()
        elif ch != 0x0D:
            cell.append(ch)
    cells.append(unquote(cell))
    return cells


def test_csv_chunked_parse_equals_line_model(tmp_path):
    """Files above a megabyte are split at line feeds and parsed on several
    threads (every LF ends a row in this dialect, quoted or not).  A 6 MB
    file of random lines (quoted cells with separators and "" escapes,
    stray quotes, CRs, empty lines, a BOM, extra cells, no final LF) gives
    the rows of the line-by-line model; a short row deep in the file is
    reported with its own line number."""
    rng = np.random.default_rng(5)
    alphabet = np.frombuffer(b'abcdefgh XYZ,;"\r' + bytes(range(0x80, 0x84)), np.uint8)

    def cell():
        k = rng.integers(0, 6)
        raw = bytes(rng.choice(alphabet, rng.integers(0, 24)))
        if k == 0:
            return b'"' + raw.replace(b'"', b'""') + b'"'  # quoted, escapes kept whole
        if k == 1:
            return b""
        raw = raw.replace(b",", b"")  # may hold stray quotes and CRs
        # a leading quote (after dropped CRs) would open a quoted cell
        return b"x" + raw if raw.lstrip(b"\r")[:1] == b'"' else raw

    lines = []
    for i in range(110000):
        if i and rng.random() < 0.01:
            lines.append(b"a,b,c,d" if rng.random() < 0.5 else b"\r")
            continue
        n = 4 + int(rng.integers(0, 3))
        lines.append(b",".join(cell() for _ in range(n)))
    lines[0] = b"h0,h1,h2,h3"
    data = b"\xef\xbb\xbf" + b"\n".join(lines)  # no final LF
    assert len(data) > (4 << 20)
    want = []
    for ln in lines:
        cells = _csv_line_model(ln, 0x2C)
        assert len(cells) >= 4 or ln == b"\r", ln
        want.append(cells[:4] if len(cells) >= 4 else None)
    # "\r" lines give one empty cell: too short, so the file with them fails
    # at the first one; the file without them parses to the model's rows
    first_short = next(i for i, w in enumerate(want) if w is None)
    p = tmp_path / "big.csv"
    p.write_bytes(data)
    with pytest.raises(capi.HBError) as e:
        hb.HostTable.from_csv(str(p))
    assert e.value.status == capi.HB_ERR_PARSE
    assert "line %d has 1 cells" % (first_short + 1) in str(e.value), str(e.value)
    keep = [ln for ln, w in zip(lines, want) if w is not None]
    p.write_bytes(b"\xef\xbb\xbf" + b"\n".join(keep))
    t = hb.HostTable.from_csv(str(p))
    assert t.n_rows == len(keep) and t.n_cols == 4
    got = _read_all(t)
    assert got == [w for w in want if w is not None]


def test_csv_errors(tmp_path):
    with pytest.raises(capi.HBError) as e:
        hb.HostTable.from_csv(str(tmp_path / "missing.csv"))
    assert e.value.status == capi.HB_ERR_IO
    p = tmp_path / "short.csv"
    p.write_bytes(b"a,b,c\nd,e\n")
    with pytest.raises(capi.HBError) as e:
        hb.HostTable.from_csv(str(p))
    assert e.value.status == capi.HB_ERR_PARSE
    p.write_bytes(b"")
    with pytest.raises(capi.HBError):
        hb.HostTable.from_csv(str(p))


# ---- partitioning -----------------------------------------------------------------

_HASH_SRC = r"""
#include <functional>
#include <iostream>
#include <string>
int main() {
  std::string s;
  while (std::getline(std::cin, s)) std::cout << std::hash<std::string>{}(s) << "\n";
}
"""


@pytest.fixture(scope="module")
def std_hash(tmp_path_factory):
    """std::hash<std::string> of the host's libstdc++, via a tiny g++ program."""
    d = tmp_path_factory.mktemp("hash")
    (d / "h.cpp").write_text(_HASH_SRC)
    subprocess.check_call(["g++", "-O1", "-o", str(d / "h"), str(d / "h.cpp")])

    def f(strings):
        out = subprocess.run([str(d / "h")], input=b"".join(s + b"\n" for s in strings),
                             stdout=subprocess.PIPE, check=True).stdout.split()
        return [int(x) for x in out]
    return f


def _bucket_model(table, col, P, prefix, hfun):
    """Bucket rule of data_mngr.h:146-206 as stated: b = hash(row[col]) % P;
    with a prefix column k and |row[k]| > 2, b = (b + (hash(row[k][:1]) + b)
    % P) % P (std::hash<int> is the identity)."""
    rows = _read_all(table)
    hs = hfun([r[col] for r in rows]) if col is not None else [0] * len(rows)
    b = [(h % P) if col is not None else 0 for h in hs]
    if prefix is not None:
        hp = hfun([r[prefix][:1] for r in rows])
        b = [(bi + (hpi + bi) % P) % P if len(r[prefix]) > 2 else bi
             for bi, hpi, r in zip(b, hp, rows)]
    return b


@pytest.mark.parametrize("P,prefix", [(1, None), (7, None), (16, None), (5, 1), (16, 2)])
def test_partition_matches_bucket_rule(std_hash, P, prefix):
    Lh, Rh, plan = H.load_fixture("dblp_acm")  # slot 0 = year (column 4) equality
    parts = hb.partition(Lh, Rh, plan, P, CP if prefix is None else prefix)
    bl = _bucket_model(Lh, 4, P, prefix, std_hash)
    br = _bucket_model(Rh, 4, P, prefix, std_hash)
    want = []
    for k in range(P):
        lr = [i for i, b in enumerate(bl) if b == k]
        rr = [i for i, b in enumerate(br) if b == k]
        if lr and rr:  # empty-side buckets are dropped
            want.append((lr, rr))
    assert [(a.tolist(), b.tolist()) for a, b in parts] == want
    assert len(parts) <= P


@pytest.mark.parametrize("P,prefix", [(1, None), (7, None), (16, 1)])
def test_partition_of_column_views_equals_csv_tables(P, prefix):
    """hb_partition_views buckets the same cells as hb_partition does for
    the CSV-loaded tables."""
    Lh, Rh, plan = H.load_fixture("dblp_acm")
    L, R = hb.ColumnTable(H.host_columns(Lh)), hb.ColumnTable(H.host_columns(Rh))
    pre = CP if prefix is None else prefix
    a = hb.partition(Lh, Rh, plan, P, pre)
    b = hb.partition(L, R, plan, P, pre)
    assert [(x.tolist(), y.tolist()) for x, y in a] == [(x.tolist(), y.tolist()) for x, y in b]


def test_partition_without_leading_equality_keeps_one_bucket(std_hash):
    Lh, Rh, plan = H.load_fixture("dblp_google_self")  # slot 0 is a JW predicate
    parts = hb.partition(Lh, Rh, plan, 8)
    assert len(parts) == 1 and len(parts[0][0]) == Lh.n_rows


def test_partitions_preserve_matches_of_equality_led_rules():
    """Hashing the first equality column keeps every equal-key pair in one
    partition, so the union of per-partition results is the whole result."""
    Lh, Rh, plan = H.load_fixture("dblp_acm")
    L, R = H.host_columns(Lh), H.host_columns(Rh)
    whole, _, _ = H.run_oracle(plan, L, R, audit=False)
    got = []
    for lr, rr in hb.partition(Lh, Rh, plan, 6):
        p, _, _ = H.run_oracle(plan, [c[lr] for c in L], [c[rr] for c in R], audit=False)
        got += [(int(lr[a]), int(rr[b])) for a, b in p]
    assert sorted(got) == [tuple(x) for x in whole.tolist()]


# ---- writer ------------------------------------------------------------------------

def test_write_pairs_lines_and_append(tmp_path):
    Lh, Rh, plan = H.load_fixture("dblp_acm")
    out = str(tmp_path / "o.csv")
    pairs = np.array([[0, 0], [1, 5], [7, 2]], np.uint32)
    hb.write_pairs(out, Lh, Rh, pairs)
    want = b"".join(Lh.cell(l, 0) + b"," + Rh.cell(r, 0) + b"\n" for l, r in pairs)
    assert open(out, "rb").read() == want
    assert want.startswith(b"id,id\n")
    hb.write_pairs(out, Lh, Rh, pairs[:1], append=True)
    assert open(out, "rb").read() == want + b"id,id\n"
    hb.write_pairs(out, Lh, Rh, pairs[1:2])  # truncates by default
    assert open(out, "rb").read() == want.split(b"\n")[1] + b"\n"
    # partition-local row numbers mapped through the partition's row lists
    hb.write_pairs(out, Lh, Rh, np.array([[1, 0]], np.uint32),
                   l_rows=np.array([3, 9]), r_rows=np.array([4]))
    assert open(out, "rb").read() == Lh.cell(9, 0) + b"," + Rh.cell(4, 0) + b"\n"
    with pytest.raises(capi.HBError):
        hb.write_pairs(out, Lh, Rh, np.array([[Lh.n_rows, 0]], np.uint32))


# ---- bound filter coefficients ---------------------------------------------------

THRESHOLDS = [0.3, 0.5, 0.7, 0.79, 0.8, 0.81, 0.85, 0.88, 0.89, 0.9, 0.95, 0.99, 0.999]


@pytest.mark.parametrize("thr", THRESHOLDS)
def test_filter_bound_never_drops_a_kept_pair(thr):
    """The device filter rejects a pair iff
         float(t) * (float(1/S) + float(1/T)) < 2 * coef[q]
    where t = 2 * (upper bound of m) >= 2 * m, q = 0 when the first bytes
    differ (p = 0) and q = 3 when only a later one of the first 4 bytes
    differs (1 <= p <= 3); otherwise it passes.  For every (m, S, T, p) with
    S <= T <= 255 whose JW is not below thr - 1e-6 (a match or an audit-band
    entry), the test above must pass at t = 2m (the left side only grows with
    t).  JW comes from the oracle's tail."""
    coef = hb.filter_coefs(thr)
    assert coef[0] >= coef[1] >= coef[2] >= coef[3]
    k = {0: np.float32(2.0) * np.float32(coef[0]), 3: np.float32(2.0) * np.float32(coef[3])}
    thr_d = float(np.float32(thr))
    rl = np.float32(1.0) / np.arange(1, 256, dtype=np.float32)  # rlen of lengths 1..255
    for S in range(1, 256):
        T = np.arange(S, 256)
        m = np.arange(0, S + 1)
        for p in range(0, min(S, 3) + 1):
            mm, TT = np.meshgrid(m, T, indexing="ij")
            keep_m = mm.ravel() >= p  # p prefix bytes match each other
            mstp = np.stack([mm.ravel(), np.full(mm.size, S), TT.ravel(),
                             np.full(mm.size, p)], axis=1)[keep_m]
            jw = O.jw_from_counts_batch(mstp)
            kept = jw > thr_d - 1e-6
            if not kept.any():
                continue
            mk, Tk = mstp[kept, 0], mstp[kept, 2]
            lhs = (2 * mk).astype(np.float32) * (rl[S - 1] + rl[Tk - 1])
            rejected = lhs < k[0 if p == 0 else 3]
            assert not rejected.any(), (thr, S, p, mstp[kept][rejected][:3])


def test_filter_bound_rejects_far_pairs():
    """The bound is useful: at 0.89, pairs of lengths 70/70 with fewer than 55
    common byte classes and a different first byte are rejected."""
    coef = hb.filter_coefs(0.89)
    lhs = np.float32(2 * 50) * (np.float32(1 / 70) + np.float32(1 / 70))
    assert lhs < np.float32(2) * np.float32(coef[0])


# ---- the executable up to the GPU --------------------------------------------------

def _exe():
    return os.path.join(H.ROOT, "hyperblocker_amd", "bin", "run_hyperblocker_exec")


def test_cli_flags_and_front_end(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    d = os.path.join(H.GOLDEN, "dblp_acm")
    base = [_exe(), "-data_l", os.path.join(d, "tableA.csv"),
            "--data_r=" + os.path.join(d, "tableB.csv"), "-rule_dir=" + os.path.join(d, "rules"),
            "-o", str(tmp_path / "out.csv"), "-n_partitions", "4", "-noread_header",
            "-scheduler=EvenSplit"]
    r = subprocess.run(base, capture_output=True, text=True)
    # plan, CSVs and partitions are done on the host before the device is needed
    assert "SerializedExecutionPlan - n_rules: 1 length: 7" in r.stdout
    assert re.search(r"Actual number of partitions: \d+", r.stdout)
    assert r.returncode != 0 and "no HIP device" in r.stderr
    r = subprocess.run(base + ["-bogus=1"], capture_output=True, text=True)
    assert r.returncode != 0 and "unknown command line flag 'bogus'" in r.stderr
    r = subprocess.run(base[:3], capture_output=True, text=True)
    assert r.returncode != 0 and "are required" in r.stderr
    r = subprocess.run(base + ["-read_header=maybe"], capture_output=True, text=True)
    assert r.returncode != 0 and "illegal value" in r.stderr


# ---- synthetic data -----------------------------------------------------------------

def test_generator_is_deterministic_and_shaped():
    a = datagen.make("c1")
    b = datagen.make("c1")
    for x, y in zip(a[0] + a[1], b[0] + b[1]):
        assert np.array_equal(x, y)
    L, R, dup, cfg = a
    assert len(L[0]) == 2600 and len(R[0]) == 2300 and len(dup) == 2200
    lens = np.char.str_len(L[1])
    assert lens.min() >= 40 and lens.max() <= 120
    for col in L + R:
        assert not any(b"," in v or b'"' in v for v in col[:500])
    years = L[4].astype(int)
    assert years.min() >= 1990 and years.max() <= 2010
