"""ctypes loader for the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
CHECK_POINT = 2147483647


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


class _Table(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int32),
                ("col_base", C.POINTER(C.c_char_p)),
                ("col_stride", C.POINTER(C.c_int64))]


class _Plan(C.Structure):
    _fields_ = [("n_rules", C.c_int32), ("length", C.c_int32),
                ("pred_index", C.POINTER(C.c_int32)),
                ("pred_type", C.c_char_p),
                ("pred_threshold", C.POINTER(C.c_float))]


class Counters(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("eq_evals", C.c_uint64),
                ("jw_evals", C.c_uint64), ("work", C.c_uint64),
                ("matches", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


AUDIT_DTYPE = np.dtype([("l", "<u4"), ("r", "<u4"), ("slot", "<i4"),
                        ("flags", "<i4"), ("jw", "<f8")])


class _Audit(C.Structure):
    _fields_ = [("l", C.c_uint32), ("r", C.c_uint32), ("slot", C.c_int32),
                ("flags", C.c_int32), ("jw", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.hbo_equal.argtypes = [C.c_char_p, C.c_char_p]
        L.hbo_equal.restype = C.c_int
        L.hbo_jaro.argtypes = [C.c_char_p, C.c_char_p]
        L.hbo_jaro.restype = C.c_float
        L.hbo_jaro_windowed.argtypes = [C.c_char_p, C.c_char_p]
        L.hbo_jaro_windowed.restype = C.c_float
        L.hbo_jw.argtypes = [C.c_char_p, C.c_char_p]
        L.hbo_jw.restype = C.c_double
        L.hbo_jw_fused.argtypes = [C.c_char_p, C.c_char_p]
        L.hbo_jw_fused.restype = C.c_double
        L.hbo_jw_work.argtypes = [C.c_char_p, C.c_char_p]
        L.hbo_jw_work.restype = C.c_uint64
        L.hbo_jw_from_counts.argtypes = [C.c_int32] * 4
        L.hbo_jw_from_counts.restype = C.c_double
        L.hbo_jw_from_counts_batch.argtypes = [C.POINTER(C.c_int32), C.c_int64,
                                               C.POINTER(C.c_double)]
        L.hbo_jw_from_counts_batch.restype = None
        L.hbo_jaro_counts.argtypes = [C.c_char_p, C.c_char_p] + [C.POINTER(C.c_int32)] * 4
        L.hbo_jaro_counts.restype = None
        L.hbo_plan_check.argtypes = [C.POINTER(_Plan)]
        L.hbo_plan_check.restype = C.c_int
        L.hbo_eval_pair.argtypes = [C.POINTER(_Plan), C.POINTER(_Table), C.c_int64,
                                    C.POINTER(_Table), C.c_int64, C.POINTER(Counters)]
        L.hbo_eval_pair.restype = C.c_int
        L.hbo_block.argtypes = [C.POINTER(_Plan), C.POINTER(_Table), C.POINTER(_Table),
                                C.c_int64, C.c_int64, C.c_int,
                                C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(Counters)]
        L.hbo_block.restype = C.c_int64
        L.hbo_audit.argtypes = [C.POINTER(_Plan), C.POINTER(_Table), C.POINTER(_Table),
                                C.c_int64, C.c_int64, C.c_int,
                                C.POINTER(C.POINTER(_Audit))]
        L.hbo_audit.restype = C.c_int64
        L.hbo_eval_pairs.argtypes = [C.POINTER(_Plan), C.POINTER(_Table), C.POINTER(_Table),
                                     C.POINTER(C.c_uint32), C.c_int64, C.c_int,
                                     C.POINTER(C.c_uint8)]
        L.hbo_eval_pairs.restype = C.c_int
        L.hbo_free.argtypes = [C.c_void_p]
        L.hbo_free.restype = None
        _lib = L
    return _lib


def _b(s):
    return s if isinstance(s, bytes) else s.encode("latin-1")


def equal(a, b):
    return bool(lib().hbo_equal(_b(a), _b(b)))


def jaro(a, b):
    return float(lib().hbo_jaro(_b(a), _b(b)))


def jaro_windowed(a, b):
    return float(lib().hbo_jaro_windowed(_b(a), _b(b)))


def jw(a, b):
    return float(lib().hbo_jw(_b(a), _b(b)))


def jw_fused(a, b):
    return float(lib().hbo_jw_fused(_b(a), _b(b)))


def jw_work(a, b):
    return int(lib().hbo_jw_work(_b(a), _b(b)))


def jw_from_counts(m, S, T, p):
    return float(lib().hbo_jw_from_counts(m, S, T, p))


def jw_from_counts_batch(mstp):
    """mstp: int32 array [n, 4] of (m, S, T, p) -> float64 [n]."""
    t = np.ascontiguousarray(np.asarray(mstp, dtype=np.int32).reshape(-1, 4))
    out = np.zeros(len(t), np.float64)
    lib().hbo_jw_from_counts_batch(t.ctypes.data_as(C.POINTER(C.c_int32)), len(t),
                                   out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def jaro_counts(a, b):
    v = [C.c_int32() for _ in range(4)]
    lib().hbo_jaro_counts(_b(a), _b(b), *[C.byref(x) for x in v])
    return tuple(int(x.value) for x in v)  # m, S, T, p


class Table:
    """Holds fixed-width NUL-padded column arrays (numpy 'S' dtype, itemsize
    >= longest cell + 1) and exposes them as an hbo_table."""

    def __init__(self, columns):
        self.columns = [np.ascontiguousarray(c) for c in columns]
        n = {len(c) for c in self.columns}
        if len(n) != 1:
            raise ValueError("ragged columns")
        self.n_rows = n.pop()
        for c in self.columns:
            if c.dtype.kind != "S":
                raise ValueError("columns must be numpy bytes arrays")
            if self.n_rows and int(np.char.str_len(c).max()) >= c.dtype.itemsize:
                raise ValueError("column itemsize must exceed the longest cell")
        nc = len(self.columns)
        self._bases = (C.c_char_p * nc)(*[C.cast(c.ctypes.data, C.c_char_p) for c in self.columns])
        self._strides = (C.c_int64 * nc)(*[c.dtype.itemsize for c in self.columns])
        self.view = _Table(self.n_rows, nc, self._bases, self._strides)


class Plan:
    def __init__(self, n_rules, pred_index, pred_type, pred_threshold):
        self.pred_index = (C.c_int32 * len(pred_index))(*pred_index)
        self.pred_type = bytes(pred_type if isinstance(pred_type, (bytes, bytearray))
                               else "".join(pred_type).encode())
        self.pred_threshold = (C.c_float * len(pred_threshold))(*pred_threshold)
        self.view = _Plan(n_rules, len(pred_index), self.pred_index,
                          C.c_char_p(self.pred_type), self.pred_threshold)

    def check(self):
        return int(lib().hbo_plan_check(C.byref(self.view)))


def block(plan, L, R, l_begin=0, l_step=1, n_threads=None):
    """Returns (pairs uint32[n,2] sorted by (l, r), counters dict)."""
    if n_threads is None:
        n_threads = os.cpu_count() or 1
    out = C.POINTER(C.c_uint32)()
    cnt = Counters()
    n = lib().hbo_block(C.byref(plan.view), C.byref(L.view), C.byref(R.view),
                        l_begin, l_step, n_threads, C.byref(out), C.byref(cnt))
    if n < 0:
        raise ValueError("oracle hbo_block failed: %d" % n)
    arr = np.ctypeslib.as_array(out, shape=(max(n, 1) * 2,))[: 2 * n].copy().reshape(n, 2)
    lib().hbo_free(out)
    return arr, cnt.as_dict()


def eval_pairs(plan, L, R, pairs, n_threads=None):
    """hbo_eval_pair of every (l, r) row of `pairs` (uint32 [n, 2]): bool [n]."""
    if n_threads is None:
        n_threads = os.cpu_count() or 1
    p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
    out = np.zeros(len(p), np.uint8)
    rc = lib().hbo_eval_pairs(C.byref(plan.view), C.byref(L.view), C.byref(R.view),
                              p.ctypes.data_as(C.POINTER(C.c_uint32)), len(p), n_threads,
                              out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc != 0:
        raise ValueError("oracle hbo_eval_pairs failed: %d" % rc)
    return out.astype(bool)


def audit(plan, L, R, l_begin=0, l_step=1, n_threads=None):
    if n_threads is None:
        n_threads = os.cpu_count() or 1
    out = C.POINTER(_Audit)()
    n = lib().hbo_audit(C.byref(plan.view), C.byref(L.view), C.byref(R.view),
                        l_begin, l_step, n_threads, C.byref(out))
    if n < 0:
        raise ValueError("oracle hbo_audit failed: %d" % n)
    buf = C.string_at(out, max(n, 0) * AUDIT_DTYPE.itemsize)
    lib().hbo_free(out)
    return np.frombuffer(buf, dtype=AUDIT_DTYPE).copy()
