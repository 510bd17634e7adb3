"""Seeded synthetic entity-resolution workloads (tables + YAML rules).

Schemas and sizes of the blocking workloads:
  c1  2,600 x 2,300 rows, 5 columns, 2,200 right rows are perturbed copies,
      one rule: year equal AND JW(title) >= 0.89 AND JW(author) >= 0.85
  c2  the same generator at 100,000 x 100,000 rows
  c3  1,000,000 x 1,000,000 rows, 8 columns, three rules

Columns (5-column schema): id (row number), title (printable ASCII without
',' and '"', lognormal length, median 70, clipped to [40, 120]), author
(20-80 chars), venue (one of 20 names), year (1990-2010).  The 8-column
schema adds city (10-40 chars), zip (6 digits) and note (0-30 chars).
Right-table copies get 5% per-character edits (substitute / insert /
delete / swap with the next character) in the text columns.

The generator itself is C++ (libhb_datagen.so) so that a million-row table
takes seconds.
"""
import ctypes as C
import math
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(PKG, "libhb_datagen.so")

ID, TEXT, VENUE, INT, DIGITS = 0, 1, 2, 3, 4


class _Col(C.Structure):
    _fields_ = [("kind", C.c_int32), ("min_len", C.c_int32), ("max_len", C.c_int32),
                ("log_mu", C.c_double), ("log_sigma", C.c_double), ("lo", C.c_int32),
                ("hi", C.c_int32), ("perturb", C.c_int32)]


def col(kind, min_len=0, max_len=0, mu=0.0, sigma=0.0, lo=0, hi=0, perturb=False):
    return (kind, min_len, max_len, mu, sigma, lo, hi, 1 if perturb else 0)


SCHEMA5 = [
    col(ID),
    col(TEXT, 40, 120, math.log(70.0), 0.35, perturb=True),   # title
    col(TEXT, 20, 80, perturb=True),                            # author
    col(VENUE),                                                 # venue
    col(INT, lo=1990, hi=2010),                                 # year
]
SCHEMA8 = SCHEMA5 + [
    col(TEXT, 10, 40, perturb=True),   # city
    col(DIGITS, 6, 6),                 # zip
    col(TEXT, 0, 30, perturb=True),    # note
]
NAMES5 = ["id", "title", "author", "venue", "year"]
NAMES8 = NAMES5 + ["city", "zip", "note"]


def rule_yaml(eq_cols, sims):
    """A rule file in the reference's format (data/dblp-acm/rules/0.yaml):
    eq_cols = [c, ...] (same column on both sides), sims = [(c, thr), ...]."""
    eq = [0, 1] + [x for c in eq_cols for x in (c, c)]
    sim = [0, 1] + [x for c, _ in sims for x in (c, c)]
    thr = ["0", "1"] + [repr(float(t)) for _, t in sims]
    return ("Preconditions:\n"
            "  Relations:\n"
            "    - [0, 0, 1, 2, 3, 4]\n"
            "    - [1, 0, 1, 2, 3, 4]\n"
            "  Equalities:\n"
            "    - [%s]\n"
            "  Sim:\n"
            "    - [%s]\n"
            "  Threshold:\n"
            "    - [%s]\n"
            "Conseq:\n"
            "  Option:\n"
            "    Equality\n"
            "  Equality:\n"
            "    [0, 1, 0, 0]\n") % (", ".join(map(str, eq)), ", ".join(map(str, sim)),
                                     ", ".join(thr))


RULES_C1 = [rule_yaml([4], [(1, 0.89), (2, 0.85)])]
RULES_C3 = [
    rule_yaml([4], [(1, 0.89), (2, 0.85)]),                 # year; title, author
    rule_yaml([3, 4], [(1, 0.80), (2, 0.90), (5, 0.85)]),   # venue, year; title, author, city
    rule_yaml([6], [(5, 0.95)]),                            # zip; city
]

CONFIGS = {
    "c1": dict(seed=1, n_l=2600, n_r=2300, n_dup=2200, schema=SCHEMA5, rules=RULES_C1,
               n_partitions=1),
    "c2": dict(seed=2, n_l=100_000, n_r=100_000, n_dup=84_615, schema=SCHEMA5,
               rules=RULES_C1, n_partitions=1),
    "c3": dict(seed=3, n_l=1_000_000, n_r=1_000_000, n_dup=846_154, schema=SCHEMA8,
               rules=RULES_C3, n_partitions=16),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError("%s missing (run python -m hyperblocker_amd.build)" % _LIB)
        L = C.CDLL(_LIB)
        L.hbg_width.argtypes = [C.POINTER(_Col)]
        L.hbg_width.restype = C.c_int64
        L.hbg_generate.argtypes = [C.c_uint64, C.c_int32, C.POINTER(_Col), C.c_int64, C.c_int64,
                                   C.c_int64, C.c_double, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_int64), C.c_int32]
        L.hbg_generate.restype = C.c_int
        L.hbg_write_csv.argtypes = [C.c_char_p, C.c_int32, C.c_int64, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_int64), C.c_char]
        L.hbg_write_csv.restype = C.c_int
        _lib = L
    return _lib


def generate(seed, n_l, n_r, n_dup, schema=SCHEMA5, edit_rate=0.05, threads=None):
    """Returns (left_columns, right_columns, dup_src): numpy 'S' arrays per
    column and, for the first n_dup right rows, the left row each copies."""
    cols = (_Col * len(schema))(*[_Col(*c) for c in schema])
    widths = [lib().hbg_width(C.byref(cols[i])) for i in range(len(schema))]
    L = [np.zeros(n_l, dtype="S%d" % w) for w in widths]
    R = [np.zeros(n_r, dtype="S%d" % w) for w in widths]
    dup = np.zeros(max(n_dup, 1), dtype=np.int64)
    pl = (C.c_void_p * len(schema))(*[a.ctypes.data for a in L])
    pr = (C.c_void_p * len(schema))(*[a.ctypes.data for a in R])
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    rc = lib().hbg_generate(seed, len(schema), cols, n_l, n_r, n_dup, edit_rate, pl, pr,
                            dup.ctypes.data_as(C.POINTER(C.c_int64)), threads)
    if rc != 0:
        raise ValueError("hbg_generate failed (%d)" % rc)
    return L, R, dup[:n_dup]


def make(name, scale=None):
    """Tables of a named workload; scale overrides (n_l, n_r) keeping the
    copy fraction."""
    cfg = dict(CONFIGS[name])
    if scale is not None:
        n_l, n_r = scale
        cfg["n_dup"] = min(n_r, int(round(n_l * cfg["n_dup"] / cfg["n_l"])))
        cfg["n_l"], cfg["n_r"] = n_l, n_r
    L, R, dup = generate(cfg["seed"], cfg["n_l"], cfg["n_r"], cfg["n_dup"], cfg["schema"])
    return L, R, dup, cfg


def write_csv(path, columns, sep=","):
    n = len(columns[0])
    bases = (C.c_void_p * len(columns))(*[a.ctypes.data for a in columns])
    widths = (C.c_int64 * len(columns))(*[a.dtype.itemsize for a in columns])
    rc = lib().hbg_write_csv(os.fsencode(path), len(columns), n, bases, widths, sep.encode())
    if rc != 0:
        raise OSError("cannot write %s" % path)


def write_dataset(directory, L, R, rules):
    """tableA.csv / tableB.csv (no header line) and rules/<k>.yaml, the layout
    of the reference's data/ directories."""
    os.makedirs(os.path.join(directory, "rules"), exist_ok=True)
    write_csv(os.path.join(directory, "tableA.csv"), L)
    write_csv(os.path.join(directory, "tableB.csv"), R)
    for k, y in enumerate(rules):
        with open(os.path.join(directory, "rules", "%d.yaml" % k), "w") as f:
            f.write(y)
    return directory
