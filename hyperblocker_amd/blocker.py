"""Python host interface of the MI355X blocking path (mirrors the reference's
HyperBlocker facade, core/hyperblocker.cuh:41-191, and its components).

All work goes through the C ABI in libhyperblocker_amd.so; nothing here
computes predicates.
"""
import ctypes as C
import itertools
import os

import numpy as np

from . import capi
from .capi import check, lib

_job_ids = itertools.count(1)


# ---- host side: rules, CSV, partitions ----------------------------------------

class HostPlan:
    """SerializedExecutionPlan built on the host
    (ExecutionPlanGenerator, core/components/execution_plan_generator.h:64-155)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.view = capi.SerializedPlan()
        check(lib().hb_host_plan_view(self._h, C.byref(self.view)))

    @classmethod
    def from_rule_dir(cls, rule_dir):
        h = C.c_void_p()
        check(lib().hb_rules_load(os.fsencode(rule_dir), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_yaml(cls, docs):
        arr = (C.c_char_p * len(docs))(*[d.encode() if isinstance(d, str) else d for d in docs])
        h = C.c_void_p()
        check(lib().hb_rules_parse(len(docs), arr, C.byref(h)))
        return cls(h.value)

    @property
    def n_rules(self):
        return self.view.n_rules

    @property
    def pred_index(self):
        return [self.view.pred_index[i] for i in range(self.view.length)]

    @property
    def pred_type(self):
        return b"".join(self.view.pred_type[i] for i in range(self.view.length)).decode()

    @property
    def pred_threshold(self):
        return [self.view.pred_threshold[i] for i in range(self.view.length)]

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hb_host_plan_free(self._h)
            self._h = None


class ArrayPlan:
    """A plan given directly as the three SerializedExecutionPlan arrays."""

    def __init__(self, n_rules, pred_index, pred_type, pred_threshold):
        n = len(pred_index)
        if not (len(pred_type) == n == len(pred_threshold)):
            raise ValueError("plan arrays differ in length")
        self._idx = (C.c_int32 * n)(*pred_index)
        pt = pred_type.encode() if isinstance(pred_type, str) else bytes(pred_type)
        self._typ = C.create_string_buffer(pt, n)
        self._thr = (C.c_float * n)(*pred_threshold)
        self.view = capi.SerializedPlan(n_rules, n, self._idx,
                                        C.cast(self._typ, C.POINTER(C.c_char)), self._thr)
        self.n_rules = n_rules
        self.pred_index = list(pred_index)
        self.pred_type = pt.decode()
        self.pred_threshold = list(pred_threshold)


class HostTable:
    """A CSV table read like rapidcsv::Document(path, LabelParams(-1, -1))
    (core/components/data_mngr.h:48-51)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        self.view = capi.TableView()
        check(lib().hb_host_table_view(self._h, C.byref(self.view)))

    @classmethod
    def from_csv(cls, path, sep=","):
        h = C.c_void_p()
        check(lib().hb_csv_load(os.fsencode(path), sep.encode()[0:1], C.byref(h)))
        return cls(h.value)

    @property
    def n_rows(self):
        return self.view.n_rows

    @property
    def n_cols(self):
        return self.view.n_cols

    def cell(self, row, col):
        v = lib().hb_host_table_cell(self._h, row, col)
        if v is None:
            raise IndexError((row, col))
        return v

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().hb_host_table_free(self._h)
            self._h = None


class ColumnTable:
    """Host table held as fixed-width numpy bytes columns ('S' dtype, itemsize
    > longest cell)."""

    def __init__(self, columns):
        self.columns = [np.ascontiguousarray(c) for c in columns]
        rows = {len(c) for c in self.columns}
        if len(rows) != 1:
            raise ValueError("ragged columns")
        self.n_rows = rows.pop()
        self.n_cols = len(self.columns)
        for c in self.columns:
            if c.dtype.kind != "S":
                raise ValueError("columns must be numpy bytes ('S') arrays")
        self._bases = (C.c_char_p * self.n_cols)(
            *[C.cast(c.ctypes.data, C.c_char_p) for c in self.columns])
        self._strides = (C.c_int64 * self.n_cols)(*[c.dtype.itemsize for c in self.columns])
        self.view = capi.TableView(self.n_rows, self.n_cols, self._bases, self._strides)

    def cell(self, row, col):
        return bytes(self.columns[col][row])


class SerializedTable:
    """A table in the reference's row-major layout (SerializedTable,
    core/data_structures/table.h:16-51, as Serialize2PinnedMemory fills it,
    :75-128): col_size[c] = longest cell + 1, col_offset[c] = sum of the
    earlier col_size (col_offset[0] = 0), aligned_tuple_size =
    (((sum col_size + 1) >> 6) << 6) + 64, cell (r, c) at
    data[r * aligned_tuple_size + col_offset[c]], zero padded.  Built from
    numpy bytes ('S') columns; `data` is a uint8 array."""

    def __init__(self, columns):
        cols = [np.ascontiguousarray(c) for c in columns]
        rows = {len(c) for c in cols}
        if len(rows) != 1:
            raise ValueError("ragged columns")
        self.n_rows = rows.pop()
        self.n_cols = len(cols)
        self.col_size = [int(np.char.str_len(c).max()) + 1 if self.n_rows else 1 for c in cols]
        self.col_offset = [int(x) for x in np.concatenate([[0], np.cumsum(self.col_size)[:-1]])]
        self.aligned_tuple_size = (((sum(self.col_size) + 1) >> 6) << 6) + 64
        ats = self.aligned_tuple_size
        self.data = np.zeros(max(self.n_rows, 1) * ats, dtype=np.uint8)
        rowview = self.data[: self.n_rows * ats].reshape(self.n_rows, ats)
        for c, off, size in zip(cols, self.col_offset, self.col_size):
            cells = np.frombuffer(c.astype("S%d" % size).tobytes(), dtype=np.uint8)
            rowview[:, off:off + size] = cells.reshape(self.n_rows, size)
        self._size = (C.c_size_t * self.n_cols)(*self.col_size)
        self._offset = (C.c_size_t * self.n_cols)(*self.col_offset)
        self.view = capi.SerializedTable(self.n_rows, self.n_cols, ats, self._size,
                                         self._offset, self.data.ctypes.data)


class DeviceBytes:
    """A buffer of device memory holding a copy of host bytes (hipMalloc +
    hipMemcpy through libamdhip64), for callers that hand the library data
    already resident on the GPU (hb_table_from_device)."""

    _hip = None

    def __init__(self, host_bytes, device=0):
        if DeviceBytes._hip is None:
            DeviceBytes._hip = C.CDLL("libamdhip64.so")
        hip = DeviceBytes._hip
        buf = np.ascontiguousarray(host_bytes).view(np.uint8)
        self.n_bytes = max(int(buf.nbytes), 1)
        self.ptr = C.c_void_p()
        if hip.hipSetDevice(device) != 0 or hip.hipMalloc(C.byref(self.ptr),
                                                          C.c_size_t(self.n_bytes)) != 0:
            raise MemoryError("hipMalloc of %d bytes failed" % self.n_bytes)
        if buf.nbytes and hip.hipMemcpy(self.ptr, C.c_void_p(buf.ctypes.data),
                                        C.c_size_t(buf.nbytes), 1) != 0:  # host to device
            raise RuntimeError("hipMemcpy to the device failed")

    def to_host(self, n_bytes=None):
        """The buffer's first n_bytes (all by default) copied to host memory."""
        n = self.n_bytes if n_bytes is None else int(n_bytes)
        out = np.zeros(max(n, 1), np.uint8)
        if n and DeviceBytes._hip.hipMemcpy(C.c_void_p(out.ctypes.data), self.ptr,
                                            C.c_size_t(n), 2) != 0:  # device to host
            raise RuntimeError("hipMemcpy from the device failed")
        return out[:n]

    def free(self):
        if self.ptr and self.ptr.value:
            DeviceBytes._hip.hipDeviceSynchronize()
            DeviceBytes._hip.hipFree(self.ptr)
        self.ptr = None


class DeviceSerializedTable:
    """A SerializedTable whose rows sit in device memory: the reference's
    d_tb_data with its col_size / col_offset (global_func.cuh:104-110)."""

    def __init__(self, table, device=0):
        self.host = table
        self.n_rows = table.n_rows
        self.buffer = DeviceBytes(table.data[: max(table.n_rows * table.aligned_tuple_size, 1)],
                                  device)
        self.view = capi.SerializedTable(table.n_rows, table.n_cols, table.aligned_tuple_size,
                                         table._size, table._offset, self.buffer.ptr.value)

    def free(self):
        self.buffer.free()


def partition(l_table, r_table, plan, n_partitions, prefix_idx=capi.CHECK_POINT):
    """DataPartitioning (core/components/data_mngr.h:81-222) of two HostTables
    or two ColumnTables.  Returns a list of (l_rows, r_rows) int64 arrays,
    empty-side partitions dropped."""
    h = C.c_void_p()
    if isinstance(l_table, ColumnTable) and isinstance(r_table, ColumnTable):
        check(lib().hb_partition_views(C.byref(l_table.view), C.byref(r_table.view),
                                       C.byref(plan.view), n_partitions, prefix_idx, C.byref(h)))
    else:
        check(lib().hb_partition(l_table._h, r_table._h, C.byref(plan.view), n_partitions,
                                 prefix_idx, C.byref(h)))
    try:
        out = []
        for k in range(lib().hb_partitions_count(h)):
            lp, rp = C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)()
            nl, nr = C.c_int64(), C.c_int64()
            check(lib().hb_partitions_get(h, k, C.byref(lp), C.byref(nl), C.byref(rp),
                                          C.byref(nr)))
            out.append((np.ctypeslib.as_array(lp, (nl.value,)).copy() if nl.value else
                        np.zeros(0, np.int64),
                        np.ctypeslib.as_array(rp, (nr.value,)).copy() if nr.value else
                        np.zeros(0, np.int64)))
        return out
    finally:
        lib().hb_partitions_free(h)


def write_pairs(path, l_table, r_table, pairs, l_rows=None, r_rows=None, append=False):
    """WriteMatch (core/components/host_reducer.h:80-90)."""
    pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
    lr = None if l_rows is None else np.ascontiguousarray(l_rows, dtype=np.int64)
    rr = None if r_rows is None else np.ascontiguousarray(r_rows, dtype=np.int64)
    check(lib().hb_write_pairs(os.fsencode(path), 1 if append else 0, l_table._h, r_table._h,
                               None if lr is None else lr.ctypes.data_as(C.POINTER(C.c_int64)),
                               None if rr is None else rr.ctypes.data_as(C.POINTER(C.c_int64)),
                               pairs.ctypes.data_as(C.POINTER(C.c_uint32)), len(pairs)))


def filter_coefs(threshold):
    out = (C.c_float * 4)()
    check(lib().hb_filter_coefs(threshold, out))
    return [out[i] for i in range(4)]


# ---- device side -----------------------------------------------------------------

class Context:
    """One GPU (hb_ctx_create)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().hb_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self):
        if self._h and self._h.value:
            check(lib().hb_ctx_destroy(self._h))
            self._h = None

    def synchronize(self):
        check(lib().hb_ctx_synchronize(self._h))

    def memory(self):
        """Device memory state (hb_ctx_memory_info), bytes."""
        m = capi.MemoryInfo()
        check(lib().hb_ctx_memory_info(self._h, C.byref(m)))
        return {k: getattr(m, k) for k, _ in m._fields_}

    @property
    def filter_shape(self):
        """(rows per lane, B rows per work item, work-groups per CU) of the
        bound-filter kernel."""
        v = [C.c_int32() for _ in range(3)]
        check(lib().hb_ctx_get_filter_shape(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    @filter_shape.setter
    def filter_shape(self, shape):
        check(lib().hb_ctx_set_filter_shape(self._h, *[int(x) for x in shape]))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, table, rows=None):
        return DeviceTable(self, table, rows)

    def from_device(self, table):
        """hb_table_from_device of a DeviceSerializedTable (rows already in
        device memory; keep it alive until jobs on the table are done)."""
        h = C.c_void_p()
        check(lib().hb_table_from_device(self._h, C.byref(table.view), C.byref(h)))
        return DeviceTable(self, table, _handle=h, _n_rows=table.n_rows)

    def upload_serialized(self, table, rows=None):
        """hb_table_upload_serialized of a SerializedTable (the reference's
        row-major handoff), or of its rows `rows` in that order
        (hb_table_upload_serialized_rows: a partition)."""
        h = C.c_void_p()
        if rows is None:
            check(lib().hb_table_upload_serialized(self._h, C.byref(table.view), C.byref(h)))
            n = table.n_rows
        else:
            rows = np.ascontiguousarray(rows, dtype=np.int64)
            check(lib().hb_table_upload_serialized_rows(
                self._h, C.byref(table.view), rows.ctypes.data_as(C.POINTER(C.c_int64)),
                len(rows), C.byref(h)))
            n = len(rows)
        return DeviceTable(self, table, _handle=h, _n_rows=n)

    def plan(self, host_plan):
        return DevicePlan(self, host_plan)

    def submit(self, plan, l_table, r_table, job_id=None):
        jid = next(_job_ids) if job_id is None else job_id
        check(lib().hb_job_submit(self._h, plan._h, l_table._h, r_table._h, jid))
        return Job(self, jid)

    def submit_batch(self, jobs, job_ids=None):
        """jobs: list of (plan, l_table, r_table).  One batch: a single launch
        sequence covers every job.  Returns the Job handles."""
        n = len(jobs)
        ids = [next(_job_ids) for _ in range(n)] if job_ids is None else list(job_ids)
        plans = (C.c_void_p * n)(*[p._h.value for p, _, _ in jobs])
        ls = (C.c_void_p * n)(*[l._h.value for _, l, _ in jobs])
        rs = (C.c_void_p * n)(*[r._h.value for _, _, r in jobs])
        idarr = (C.c_uint64 * n)(*ids)
        check(lib().hb_batch_submit(self._h, n, plans, ls, rs, idarr))
        return [Job(self, i) for i in ids]

    def batch_graph(self, jobs):
        """jobs: list of (plan, l_table, r_table), run repeatedly as one
        captured batch (hb_batch_graph_create).  Returns a BatchGraph."""
        n = len(jobs)
        plans = (C.c_void_p * n)(*[p._h.value for p, _, _ in jobs])
        ls = (C.c_void_p * n)(*[l._h.value for _, l, _ in jobs])
        rs = (C.c_void_p * n)(*[r._h.value for _, _, r in jobs])
        h = C.c_void_p()
        check(lib().hb_batch_graph_create(self._h, n, plans, ls, rs, C.byref(h)))
        return BatchGraph(self, h, n)

    def batch_results(self, job):
        """All results of the batch `job` ran in (hb_batch_results): (job ids,
        pairs uint32[n, 2], pair offsets uint64[n_jobs + 1], audit entries,
        audit offsets), copied out.  Job k's pairs are
        pairs[pair_off[k]:pair_off[k + 1]]."""
        n = C.c_uint32()
        ids, po, ao = (C.POINTER(C.c_uint64)() for _ in range(3))
        p = C.POINTER(C.c_uint32)()
        a = C.POINTER(capi.AuditEntry)()
        check(lib().hb_batch_results(self._h, job.id, C.byref(n), C.byref(ids), C.byref(p),
                                     C.byref(po), C.byref(a), C.byref(ao)))
        nj = n.value
        ids_a = np.ctypeslib.as_array(ids, (nj,)).copy()
        po_a = np.ctypeslib.as_array(po, (nj + 1,)).copy()
        ao_a = np.ctypeslib.as_array(ao, (nj + 1,)).copy()
        n_pairs, n_audit = int(po_a[-1]), int(ao_a[-1])
        pairs = (np.ctypeslib.as_array(p, (2 * n_pairs,)).reshape(-1, 2).copy()
                 if n_pairs else np.zeros((0, 2), np.uint32))
        audit = (np.frombuffer(C.string_at(a, n_audit * capi.AUDIT_DTYPE.itemsize),
                               dtype=capi.AUDIT_DTYPE).copy()
                 if n_audit else np.zeros(0, capi.AUDIT_DTYPE))
        return ids_a, pairs, po_a, audit, ao_a

    def release_jobs(self, jobs):
        """hb_job_release of every job, in one call."""
        ids = (C.c_uint64 * len(jobs))(*[j.id for j in jobs])
        check(lib().hb_jobs_release(self._h, len(jobs), ids))

    def autotune(self, jobs, candidates=None):
        """Picks the bound filter's launch shape for these jobs (list of
        (plan, l_table, r_table)) and makes it the context's; returns
        (shape, filter ms)."""
        n = len(jobs)
        plans = (C.c_void_p * n)(*[p._h.value for p, _, _ in jobs])
        ls = (C.c_void_p * n)(*[l._h.value for _, l, _ in jobs])
        rs = (C.c_void_p * n)(*[r._h.value for _, _, r in jobs])
        cand = None
        nc = 0
        if candidates:
            flat = [int(x) for c in candidates for x in c]
            cand = (C.c_int32 * len(flat))(*flat)
            nc = len(candidates)
        best = (C.c_int32 * 3)()
        ms = C.c_double()
        check(lib().hb_autotune_filter_shape(self._h, n, plans, ls, rs, cand, nc, best,
                                             C.byref(ms)))
        return (best[0], best[1], best[2]), ms.value

    def block_noalloc(self, plan, l_table, r_table, pairs_buf, pair_cap, audit_buf, audit_cap):
        """hb_block_noalloc: the job of block() on prepared tables, writing
        into device buffers (DeviceBytes); returns (n_pairs, n_audit).
        Raises HBError (status HB_ERR_CAPACITY) when the outputs do not fit,
        with the counts in .counts."""
        n, na = C.c_uint64(), C.c_uint64()
        rc = lib().hb_block_noalloc(self._h, plan._h, l_table._h, r_table._h,
                                    pairs_buf.ptr if pairs_buf else None, pair_cap,
                                    audit_buf.ptr if audit_buf else None, audit_cap,
                                    C.byref(n), C.byref(na))
        if rc != capi.HB_OK:
            e = capi.HBError(rc, lib().hb_last_error().decode(errors="replace"))
            e.counts = (n.value, na.value)
            raise e
        return n.value, na.value

    def block(self, plan, l_table, r_table):
        """Synchronous blocking of l_table x r_table: (pairs uint32[n,2] in no
        particular order, audit structured array, stats dict)."""
        job = self.submit(plan, l_table, r_table)
        try:
            return job.results()
        finally:
            job.release()


class Server:
    """Serving loop on a context (hb_server_*): jobs are queued and admitted
    into batches as batch slots and working memory free up."""

    def __init__(self, ctx, max_batches_in_flight=2, max_jobs_per_batch=256, memory_budget=0):
        self.ctx = ctx
        cfg = capi.ServerConfig(max_batches_in_flight, max_jobs_per_batch, memory_budget)
        h = C.c_void_p()
        check(lib().hb_server_create(ctx._h, C.byref(cfg), C.byref(h)))
        self._h = h

    def submit(self, plan, l_table, r_table, job_id=None):
        jid = next(_job_ids) if job_id is None else job_id
        check(lib().hb_server_submit(self._h, plan._h, l_table._h, r_table._h, jid))
        return Job(self.ctx, jid)

    def info(self):
        v = [C.c_uint64() for _ in range(3)]
        check(lib().hb_server_info(self._h, *[C.byref(x) for x in v]))
        return {"batches": v[0].value, "jobs": v[1].value, "queued": v[2].value}

    def close(self):
        if self._h and self._h.value:
            check(lib().hb_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceTable:
    def __init__(self, ctx, table, rows=None, _handle=None, _n_rows=None):
        self.ctx = ctx
        self.host = table  # keeps the source buffers alive for the caller
        if _handle is not None:
            self._h = _handle
            self.n_rows = _n_rows
            return
        h = C.c_void_p()
        if rows is None:
            check(lib().hb_table_upload(ctx._h, C.byref(table.view), C.byref(h)))
            self.n_rows = table.view.n_rows
        else:
            rows = np.ascontiguousarray(rows, dtype=np.int64)
            check(lib().hb_table_upload_rows(ctx._h, C.byref(table.view),
                                             rows.ctypes.data_as(C.POINTER(C.c_int64)),
                                             len(rows), C.byref(h)))
            self.n_rows = len(rows)
        self._h = h

    def slice(self, begin, end):
        """Rows [begin, end) as a table sharing this one's device data."""
        h = C.c_void_p()
        check(lib().hb_table_slice(self.ctx._h, self._h, int(begin), int(end), C.byref(h)))
        return DeviceTable(self.ctx, self.host, _handle=h, _n_rows=int(end) - int(begin))

    def release(self):
        if self._h and self._h.value and self.ctx._h:
            check(lib().hb_table_release(self.ctx._h, self._h))
        self._h = None


class DevicePlan:
    def __init__(self, ctx, host_plan):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().hb_plan_create(ctx._h, C.byref(host_plan.view), C.byref(h)))
        self._h = h

    def release(self):
        if self._h and self._h.value and self.ctx._h:
            check(lib().hb_plan_release(self.ctx._h, self._h))
        self._h = None


class BatchGraph:
    """A job set captured as one HIP graph (hb_batch_graph_*); each run
    replays it as a batch of new jobs."""

    def __init__(self, ctx, handle, n_jobs):
        self.ctx = ctx
        self._h = handle
        self.n_jobs = n_jobs

    def run(self, job_ids=None):
        """Replays the graph; returns the runs' Job handles (one per job, in
        the order the graph was created with)."""
        ids = ([next(_job_ids) for _ in range(self.n_jobs)] if job_ids is None
               else list(job_ids))
        check(lib().hb_batch_graph_run(self._h, (C.c_uint64 * self.n_jobs)(*ids)))
        return [Job(self.ctx, i) for i in ids]

    def release(self):
        if self._h:
            check(lib().hb_batch_graph_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Job:
    def __init__(self, ctx, job_id):
        self.ctx = ctx
        self.id = job_id

    def done(self):
        return check(lib().hb_job_query(self.ctx._h, self.id)) == capi.HB_OK

    def wait(self):
        check(lib().hb_job_wait(self.ctx._h, self.id))

    def stats(self):
        s = capi.JobStats()
        check(lib().hb_job_stats_get(self.ctx._h, self.id, C.byref(s)))
        return s.as_dict()

    def results(self):
        p = C.POINTER(C.c_uint32)()
        a = C.POINTER(capi.AuditEntry)()
        n, na = C.c_uint64(), C.c_uint64()
        check(lib().hb_job_results(self.ctx._h, self.id, C.byref(p), C.byref(n), C.byref(a),
                                   C.byref(na)))
        pairs = (np.ctypeslib.as_array(p, (n.value * 2,)).reshape(-1, 2).copy()
                 if n.value else np.zeros((0, 2), np.uint32))
        audit = (np.frombuffer(C.string_at(a, na.value * capi.AUDIT_DTYPE.itemsize),
                               dtype=capi.AUDIT_DTYPE).copy()
                 if na.value else np.zeros(0, capi.AUDIT_DTYPE))
        return pairs, audit, self.stats()

    def release(self):
        check(lib().hb_job_release(self.ctx._h, self.id))


def sort_pairs(pairs):
    """(l, r) pairs sorted lexicographically."""
    pairs = np.asarray(pairs).reshape(-1, 2)
    if len(pairs) == 0:
        return pairs.astype(np.uint32)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order].astype(np.uint32)
