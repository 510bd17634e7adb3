"""ctypes bindings of include/hyperblocker_amd.h.

Loads the in-tree libhyperblocker_amd.so and fails loudly if it is missing:
there is no Python or CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhyperblocker_amd.so")

HB_OK = 0
HB_NOT_READY = 1
HB_ERR_INVALID_ARG = -1
HB_ERR_PLAN = -2
HB_ERR_OOM = -3
HB_ERR_CAPACITY = -4
HB_ERR_HIP = -5
HB_ERR_NOT_FOUND = -6
HB_ERR_IO = -7
HB_ERR_PARSE = -8
HB_ERR_NO_DEVICE = -9
CHECK_POINT = 2147483647

# Every symbol include/hyperblocker_amd.h declares.
EXPORTS = [
    "hb_ctx_create", "hb_ctx_destroy", "hb_ctx_synchronize", "hb_ctx_memory_info",
    "hb_ctx_set_filter_shape",
    "hb_ctx_get_filter_shape", "hb_last_error", "hb_version",
    "hb_table_upload", "hb_table_upload_serialized", "hb_table_upload_serialized_rows", "hb_table_from_device", "hb_table_upload_rows",
    "hb_table_slice", "hb_table_info", "hb_table_release",
    "hb_plan_create", "hb_plan_release",
    "hb_job_submit", "hb_batch_submit", "hb_batch_graph_create", "hb_batch_graph_run",
    "hb_batch_graph_destroy", "hb_job_query", "hb_job_wait", "hb_job_results",
    "hb_job_stats_get", "hb_job_release", "hb_batch_results", "hb_jobs_release", "hb_block", "hb_free",
    "hb_block_noalloc", "hb_autotune_filter_shape",
    "hb_server_create", "hb_server_submit", "hb_server_info", "hb_server_destroy",
    "hb_rules_load", "hb_rules_parse", "hb_host_plan_view", "hb_host_plan_free",
    "hb_csv_load", "hb_host_table_view", "hb_host_table_cell", "hb_host_table_free",
    "hb_partition", "hb_partition_views", "hb_partitions_count", "hb_partitions_get", "hb_partitions_free",
    "hb_write_pairs", "hb_filter_coefs", "hb_jw_from_counts",
]


class HBError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("hyperblocker_amd error %d: %s" % (status, msg))
        self.status = status


class TableView(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int32),
                ("col_base", C.POINTER(C.c_char_p)), ("col_stride", C.POINTER(C.c_int64))]


class SerializedTable(C.Structure):
    _fields_ = [("n_rows", C.c_size_t), ("n_cols", C.c_size_t),
                ("aligned_tuple_size", C.c_size_t), ("col_size", C.POINTER(C.c_size_t)),
                ("col_offset", C.POINTER(C.c_size_t)), ("data", C.c_void_p)]


class SerializedPlan(C.Structure):
    _fields_ = [("n_rules", C.c_int32), ("length", C.c_int32),
                ("pred_index", C.POINTER(C.c_int32)), ("pred_type", C.POINTER(C.c_char)),
                ("pred_threshold", C.POINTER(C.c_float))]


class AuditEntry(C.Structure):
    _fields_ = [("l", C.c_uint32), ("r", C.c_uint32), ("slot", C.c_int32),
                ("flags", C.c_int32), ("jw", C.c_double)]


AUDIT_DTYPE = np.dtype([("l", "<u4"), ("r", "<u4"), ("slot", "<i4"), ("flags", "<i4"),
                        ("jw", "<f8")])


class ServerConfig(C.Structure):
    _fields_ = [("max_batches_in_flight", C.c_int32), ("max_jobs_per_batch", C.c_int32),
                ("memory_budget", C.c_uint64)]


class MemoryInfo(C.Structure):
    _fields_ = [("device_free", C.c_uint64), ("device_total", C.c_uint64),
                ("pool_live", C.c_uint64), ("pool_cached", C.c_uint64),
                ("pinned_cached", C.c_uint64)]


class JobStats(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("candidates", C.c_uint64), ("filtered", C.c_uint64),
                ("jw_exact", C.c_uint64), ("matches", C.c_uint64), ("audit", C.c_uint64),
                ("device_ms", C.c_double), ("join_ms", C.c_double), ("filter_ms", C.c_double),
                ("exact_ms", C.c_double), ("retries", C.c_int32), ("n_rules", C.c_int32),
                ("queue_ms", C.c_double), ("latency_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
VP = C.c_void_p
PVP = C.POINTER(C.c_void_p)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("hyperblocker_amd native library not built: %s is missing "
                          "(run python -m hyperblocker_amd.build)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    st = C.c_int32
    sig = {
        "hb_ctx_create": ([C.c_int32, PVP], st),
        "hb_ctx_destroy": ([VP], st),
        "hb_ctx_synchronize": ([VP], st),
        "hb_ctx_memory_info": ([VP, C.POINTER(MemoryInfo)], st),
        "hb_ctx_set_filter_shape": ([VP, C.c_int32, C.c_int32, C.c_int32], st),
        "hb_ctx_get_filter_shape": ([VP] + [C.POINTER(C.c_int32)] * 3, st),
        "hb_last_error": ([], C.c_char_p),
        "hb_version": ([], C.c_char_p),
        "hb_table_upload": ([VP, C.POINTER(TableView), PVP], st),
        "hb_table_upload_serialized": ([VP, C.POINTER(SerializedTable), PVP], st),
        "hb_table_from_device": ([VP, C.POINTER(SerializedTable), PVP], st),
        "hb_table_upload_serialized_rows": ([VP, C.POINTER(SerializedTable), C.POINTER(C.c_int64),
                                             C.c_int64, PVP], st),
        "hb_table_upload_rows": ([VP, C.POINTER(TableView), C.POINTER(C.c_int64), C.c_int64,
                                  PVP], st),
        "hb_table_slice": ([VP, VP, C.c_int64, C.c_int64, PVP], st),
        "hb_table_info": ([VP, C.POINTER(C.c_int64), C.POINTER(C.c_int32)], st),
        "hb_table_release": ([VP, VP], st),
        "hb_plan_create": ([VP, C.POINTER(SerializedPlan), PVP], st),
        "hb_plan_release": ([VP, VP], st),
        "hb_block_noalloc": ([VP, VP, VP, VP, VP, C.c_uint64, VP, C.c_uint64,
                              C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], st),
        "hb_job_submit": ([VP, VP, VP, VP, C.c_uint64], st),
        "hb_batch_submit": ([VP, C.c_uint32, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP),
                             C.POINTER(C.c_uint64)], st),
        "hb_batch_graph_create": ([VP, C.c_uint32, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP),
                                   PVP], st),
        "hb_batch_graph_run": ([VP, C.POINTER(C.c_uint64)], st),
        "hb_batch_graph_destroy": ([VP], st),
        "hb_job_query": ([VP, C.c_uint64], st),
        "hb_job_wait": ([VP, C.c_uint64], st),
        "hb_job_results": ([VP, C.c_uint64, C.POINTER(C.POINTER(C.c_uint32)),
                            C.POINTER(C.c_uint64), C.POINTER(C.POINTER(AuditEntry)),
                            C.POINTER(C.c_uint64)], st),
        "hb_job_stats_get": ([VP, C.c_uint64, C.POINTER(JobStats)], st),
        "hb_job_release": ([VP, C.c_uint64], st),
        "hb_batch_results": ([VP, C.c_uint64, C.POINTER(C.c_uint32),
                              C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.POINTER(C.c_uint32)),
                              C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.POINTER(AuditEntry)),
                              C.POINTER(C.POINTER(C.c_uint64))], st),
        "hb_jobs_release": ([VP, C.c_uint32, C.POINTER(C.c_uint64)], st),
        "hb_block": ([VP, VP, VP, VP, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_uint64),
                      C.POINTER(C.POINTER(AuditEntry)), C.POINTER(C.c_uint64),
                      C.POINTER(JobStats)], st),
        "hb_autotune_filter_shape": ([VP, C.c_uint32, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP),
                                      C.POINTER(C.c_int32), C.c_uint32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_double)], st),
        "hb_server_create": ([VP, C.POINTER(ServerConfig), PVP], st),
        "hb_server_submit": ([VP, VP, VP, VP, C.c_uint64], st),
        "hb_server_info": ([VP] + [C.POINTER(C.c_uint64)] * 3, st),
        "hb_server_destroy": ([VP], st),
        "hb_free": ([VP], None),
        "hb_rules_load": ([C.c_char_p, PVP], st),
        "hb_rules_parse": ([C.c_int32, C.POINTER(C.c_char_p), PVP], st),
        "hb_host_plan_view": ([VP, C.POINTER(SerializedPlan)], st),
        "hb_host_plan_free": ([VP], None),
        "hb_csv_load": ([C.c_char_p, C.c_char, PVP], st),
        "hb_host_table_view": ([VP, C.POINTER(TableView)], st),
        "hb_host_table_cell": ([VP, C.c_int64, C.c_int32], C.c_char_p),
        "hb_host_table_free": ([VP], None),
        "hb_partition": ([VP, VP, C.POINTER(SerializedPlan), C.c_uint64, C.c_int64, PVP], st),
        "hb_partition_views": ([C.POINTER(TableView), C.POINTER(TableView),
                                C.POINTER(SerializedPlan), C.c_uint64, C.c_int64, PVP], st),
        "hb_partitions_count": ([VP], C.c_int64),
        "hb_partitions_get": ([VP, C.c_int64, C.POINTER(C.POINTER(C.c_int64)),
                               C.POINTER(C.c_int64), C.POINTER(C.POINTER(C.c_int64)),
                               C.POINTER(C.c_int64)], st),
        "hb_partitions_free": ([VP], None),
        "hb_write_pairs": ([C.c_char_p, C.c_int32, VP, VP, C.POINTER(C.c_int64),
                            C.POINTER(C.c_int64), C.POINTER(C.c_uint32), C.c_uint64], st),
        "hb_filter_coefs": ([C.c_float, C.POINTER(C.c_float)], st),
        "hb_jw_from_counts": ([VP, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)], st),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def last_error():
    return lib().hb_last_error().decode("utf-8", "replace")


def check(status):
    if status not in (HB_OK, HB_NOT_READY):
        raise HBError(status, last_error())
    return status
