/*
 * hyperblocker_amd.h -- C ABI of the MI355X (gfx950) rule-based blocking path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an hb_status (0 = HB_OK).  No call truncates output silently: a
 * result that does not fit grows its buffer, and an input the reference
 * would misread is rejected with an error.
 *
 * Reference interfaces replaced (paths relative to the HyperBlocker tree):
 *   - table layout      core/data_structures/table.h:16-51 (SerializedTable)
 *   - execution plan    core/components/execution_plan_generator.h:25-30
 *                       (SerializedExecutionPlan) and :64-155
 *                       (ExecutionPlanGenerator, GetExecutionPlan)
 *   - rule files        core/common/yaml_config.h:19-57 (convert<Rule>)
 *   - CSV load          core/components/data_mngr.h:46-79 (DataMngr)
 *   - partitioning      core/components/data_mngr.h:81-222 (DataPartitioning)
 *   - submit/complete   core/components/host_producer.h:109-215 (AsyncSubmit)
 *                       core/components/host_reducer.h:32-68 (Run: polling)
 *   - kernel            core/gpu/global_func.cuh:104-185 (blocking_kernel)
 *   - output            core/components/host_reducer.h:80-90 (WriteMatch)
 */
#ifndef HYPERBLOCKER_AMD_H_
#define HYPERBLOCKER_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t hb_status;
#define HB_OK 0
#define HB_NOT_READY 1        /* job still running (hb_job_query) */
#define HB_ERR_INVALID_ARG -1 /* null/out-of-range argument */
#define HB_ERR_PLAN -2        /* malformed plan or rule */
#define HB_ERR_OOM -3         /* device or pinned allocation failed */
#define HB_ERR_CAPACITY -4    /* a configured size limit was exceeded */
#define HB_ERR_HIP -5         /* HIP runtime error */
#define HB_ERR_NOT_FOUND -6   /* unknown job id */
#define HB_ERR_IO -7          /* file could not be read or written */
#define HB_ERR_PARSE -8       /* CSV / YAML / flag syntax error */
#define HB_ERR_NO_DEVICE -9   /* no HIP device */

/* core/common/types.h:8-13 */
#define HB_EQUALITIES 'e'
#define HB_SIM 's'
#define HB_CHECK_POINT 2147483647
#define HB_CHECK_POINT_CHAR ' '

typedef struct hb_ctx hb_ctx;
typedef struct hb_table hb_table;
typedef struct hb_plan hb_plan;
typedef struct hb_host_table hb_host_table;
typedef struct hb_host_plan hb_host_plan;
typedef struct hb_partitions hb_partitions;
typedef struct hb_server hb_server;

/* Column-addressed view of a host table.  Cell (r, c) is the byte string at
 * col_base[c] + r * col_stride[c], ending at the first NUL or after
 * col_stride[c] bytes, whichever comes first. */
typedef struct {
  int64_t n_rows;
  int32_t n_cols;
  const char *const *col_base;
  const int64_t *col_stride;
} hb_table_view;

/* Same fields as SerializedTable (core/data_structures/table.h:16-51):
 * cell (r, c) at data + r * aligned_tuple_size + col_offset[c], col_size[c]
 * bytes, NUL padded. */
typedef struct {
  size_t n_rows;
  size_t n_cols;
  size_t aligned_tuple_size;
  const size_t *col_size;
  const size_t *col_offset;
  const char *data;
} hb_serialized_table;

/* Same fields as SerializedExecutionPlan
 * (core/components/execution_plan_generator.h:25-30).  Per rule: slot pairs
 * (left column, right column) of type 'e' or 's', then one checkpoint slot
 * (HB_CHECK_POINT, ' ', 1).  A similarity predicate uses the threshold of its
 * left slot, compared as double (core/gpu/global_func.cuh:139-145). */
typedef struct {
  int32_t n_rules;
  int32_t length;
  const int32_t *pred_index;
  const char *pred_type;
  const float *pred_threshold;
} hb_serialized_plan;

/* One similarity evaluation whose Jaro-Winkler value lies within 1e-6 of its
 * threshold.  slot = plan slot of the predicate's left column.
 * flags bit0: a fused multiply-add in the Winkler step would decide the
 * other way; bit1: the shorter string has >= 257 bytes (the reference's
 * 256-entry match array is exceeded there). */
typedef struct {
  uint32_t l, r;
  int32_t slot;
  int32_t flags;
  double jw;
} hb_audit_entry;

typedef struct {
  uint64_t pairs;      /* rule-evaluated pairs: n_l * n_r */
  uint64_t candidates; /* pairs passing a rule's equalities, summed over rules */
  uint64_t filtered;   /* candidates the bound filter sent to exact JW */
  uint64_t jw_exact;   /* exact JW evaluations */
  uint64_t matches;    /* emitted pairs */
  uint64_t audit;      /* audit entries */
  double device_ms;    /* first kernel start to last kernel end */
  double join_ms;      /* equality join construction, summed over rules */
  double filter_ms;    /* equality + bound filter stage, summed over rules */
  double exact_ms;     /* exact Jaro-Winkler stages, summed over rules */
  int32_t retries;     /* capacity regrowth reruns */
  int32_t n_rules;
  double queue_ms;     /* submit to admission into a batch (served jobs) */
  double latency_ms;   /* submit to completion of its batch */
} hb_job_stats;

/* ---- context -------------------------------------------------------------
 * One context per GPU (hipSetDevice(device) is made current by every call
 * that touches the device).  Thread-safe: calls may come from any thread. */
hb_status hb_ctx_create(int32_t device, hb_ctx **out);
hb_status hb_ctx_destroy(hb_ctx *ctx);
/* Waits until all work of every stream on the context's device is done. */
/* Device memory as the library sees it: the device's free and total bytes,
 * bytes the library holds from the driver (live) and, of those, bytes kept
 * in its cache for reuse, and pinned host bytes in its cache.  No
 * counterpart in the reference (which never frees device memory). */
typedef struct hb_memory_info {
  uint64_t device_free, device_total;
  uint64_t pool_live, pool_cached;
  uint64_t pinned_cached;
} hb_memory_info;
hb_status hb_ctx_memory_info(hb_ctx *ctx, hb_memory_info *info);

hb_status hb_ctx_synchronize(hb_ctx *ctx);
/* Launch shape of the bound-filter kernel on this context: rows of the
 * larger side per lane (1, 2 or 4), rows of the other side per work item
 * (a multiple of 64, at most 65536) and work-groups per compute unit
 * (1..64); (2, 192, 16) by default, or HB_FILTER_SHAPE="r,t,g" at context
 * creation.  Applies to jobs submitted afterwards; results do not depend on
 * it.  No counterpart in the reference (its launch is fixed at 64 x 64
 * threads, host_producer.h:114-115); used by the autotuner. */
hb_status hb_ctx_set_filter_shape(hb_ctx *ctx, int32_t rows_per_lane, int32_t tile_b,
                                  int32_t groups_per_cu);
hb_status hb_ctx_get_filter_shape(const hb_ctx *ctx, int32_t *rows_per_lane, int32_t *tile_b,
                                  int32_t *groups_per_cu);
/* Text of the last error on this thread (never NULL). */
const char *hb_last_error(void);
/* Library build id, e.g. "hyperblocker_amd 0.2 gfx950". */
const char *hb_version(void);

/* ---- tables --------------------------------------------------------------
 * Uploading copies the cells into library-owned pinned staging, starts the
 * host-to-device copy on the context's copy stream and returns; the caller
 * may reuse its buffers immediately.  Jobs that use the table wait for the
 * copy on the device.  A table may be shared by any number of jobs and
 * stays resident until released (releasing waits for those jobs). */
hb_status hb_table_upload(hb_ctx *ctx, const hb_table_view *view, hb_table **out);
/* The partitioner's pinned SerializedTable (core/data_structures/table.h:16-51,
 * filled by Serialize2PinnedMemory, :75-128), as AsyncSubmit hands it over
 * (host_producer.h:137-152): the rows are copied as they are (one parallel
 * copy into staging, one host-to-device copy) and unpacked into columns on
 * the device.  Cell (r, c) ends at its first NUL or after col_size[c] bytes. */
hb_status hb_table_upload_serialized(hb_ctx *ctx, const hb_serialized_table *t,
                                     hb_table **out);
/* The same for rows[0..n_rows) of the SerializedTable, in that order: a
 * partition's re-serialisation (DataPartitioning, data_mngr.h:209-215) and
 * its host-to-device copy (AsyncSubmit, host_producer.h:137-152) in one
 * call, one row copy per row. */
hb_status hb_table_upload_serialized_rows(hb_ctx *ctx, const hb_serialized_table *t,
                                          const int64_t *rows, int64_t n_rows, hb_table **out);
/* The reference's device-side table as blocking_kernel receives it
 * (d_tb_data with its col_size / col_offset, core/gpu/global_func.cuh:104-110,
 * after AsyncSubmit's host-to-device copy, host_producer.h:137-152): `data`
 * is a device pointer to n_rows rows of aligned_tuple_size bytes on this
 * context's device; col_size and col_offset are host arrays.  The cells are
 * unpacked into a new table on the device (no host copy, no host pass over
 * the rows).  The call returns before the device has read `data`: keep it
 * valid and unchanged until a job on the table has completed (or
 * hb_ctx_synchronize has returned). */
hb_status hb_table_from_device(hb_ctx *ctx, const hb_serialized_table *t, hb_table **out);
/* Uploads rows[0..n_rows) of the view, in that order, as a new table: the
 * per-partition upload of the streamed form (the reference re-serialises
 * each partition, core/components/data_mngr.h:209-215, and copies it in
 * AsyncSubmit, host_producer.h:137-152). */
hb_status hb_table_upload_rows(hb_ctx *ctx, const hb_table_view *view, const int64_t *rows,
                               int64_t n_rows, hb_table **out);
/* A table made of rows [row_begin, row_end) of `parent`, sharing its device
 * data and per-column structures (no copy; join layouts are built per
 * slice).  Used for many small jobs over a resident table (batched and
 * served forms); no counterpart in the reference, which copies every
 * partition (data_mngr.h:209-215).  The parent cannot be released while
 * slices of it are alive. */
hb_status hb_table_slice(hb_ctx *ctx, hb_table *parent, int64_t row_begin, int64_t row_end,
                         hb_table **out);
hb_status hb_table_info(const hb_table *t, int64_t *n_rows, int32_t *n_cols);
hb_status hb_table_release(hb_ctx *ctx, hb_table *t);

/* ---- plans ---------------------------------------------------------------
 * Validates and copies the plan: rejects an empty rule, a missing trailing
 * checkpoint, an unknown predicate type, mismatched slot types and negative
 * column indices.  Column indices are checked against the tables at submit. */
hb_status hb_plan_create(hb_ctx *ctx, const hb_serialized_plan *plan, hb_plan **out);
hb_status hb_plan_release(hb_ctx *ctx, hb_plan *plan);

/* ---- jobs ----------------------------------------------------------------
 * hb_job_submit enqueues the blocking of every pair (l, r) of L x R under the
 * plan (AsyncSubmit + blocking_kernel) and returns without waiting.  job_id
 * is chosen by the caller and must be unused.  Results are library-owned
 * until hb_job_release. */
hb_status hb_job_submit(hb_ctx *ctx, const hb_plan *plan, const hb_table *l,
                        const hb_table *r, uint64_t job_id);
/* Batched form: n_jobs jobs (plans[i], l[i], r[i], id job_ids[i]) run as one
 * batch, i.e. one sequence of kernel launches covering all of their rules
 * (no counterpart in the reference, which launches one kernel per
 * partition, host_producer.h:195-197).  Each job is then queried, read and
 * released by its id like a job from hb_job_submit; results equal those of
 * the same job submitted alone.  All jobs of a batch complete together. */
hb_status hb_batch_submit(hb_ctx *ctx, uint32_t n_jobs, const hb_plan *const *plans,
                          const hb_table *const *l, const hb_table *const *r,
                          const uint64_t *job_ids);
/* Captured batch: a job set run repeatedly (no counterpart in the
 * reference).  hb_batch_graph_create prepares n_jobs jobs as for
 * hb_batch_submit (per-table structures, descriptors, buffers sized from
 * the plans' hints) and captures the batch's launch sequence into a HIP
 * graph; it runs nothing.  Each hb_batch_graph_run replays the graph once
 * as a batch whose jobs get the ids job_ids[0..n_jobs) (unused ids, read,
 * waited for and released like those of hb_batch_submit; results equal the
 * jobs submitted alone).  A run first completes the previous run of the
 * graph (its buffers are the graph's); a run whose lists overflow reruns
 * with larger ones like any batch, and the next run captures the graph
 * again with those sizes.  Graph runs report no join/filter/exact stage
 * times (0), only device_ms.  The plans and tables must outlive the graph.
 * hb_batch_graph_destroy completes the last run and frees the graph. */
typedef struct hb_batch_graph hb_batch_graph;
hb_status hb_batch_graph_create(hb_ctx *ctx, uint32_t n_jobs, const hb_plan *const *plans,
                                const hb_table *const *l, const hb_table *const *r,
                                hb_batch_graph **out);
hb_status hb_batch_graph_run(hb_batch_graph *g, const uint64_t *job_ids);
hb_status hb_batch_graph_destroy(hb_batch_graph *g);
hb_status hb_job_query(hb_ctx *ctx, uint64_t job_id); /* HB_OK or HB_NOT_READY */
hb_status hb_job_wait(hb_ctx *ctx, uint64_t job_id);
/* (l_row, r_row) pairs as 2*n_pairs uint32, each matching pair exactly once,
 * in unspecified order (the reference's order is that of its atomicAdd).
 * Rows index the job's L and R tables. */
hb_status hb_job_results(hb_ctx *ctx, uint64_t job_id, const uint32_t **pairs,
                         uint64_t *n_pairs, const hb_audit_entry **audit,
                         uint64_t *n_audit);
hb_status hb_job_stats_get(hb_ctx *ctx, uint64_t job_id, hb_job_stats *out);
hb_status hb_job_release(hb_ctx *ctx, uint64_t job_id);

/* Results of the whole batch that job_id ran in, in one block (no
 * counterpart in the reference; for callers of hb_batch_submit that read
 * thousands of jobs): job k of the batch has id job_ids[k], its pairs are
 * pairs[2*pair_off[k] .. 2*pair_off[k+1]) and its audit entries
 * audit[audit_off[k] .. audit_off[k+1]), as hb_job_results gives them.
 * Waits for the batch.  The arrays are library-owned and valid while job_id
 * is not released. */
hb_status hb_batch_results(hb_ctx *ctx, uint64_t job_id, uint32_t *n_jobs,
                           const uint64_t **job_ids, const uint32_t **pairs,
                           const uint64_t **pair_off, const hb_audit_entry **audit,
                           const uint64_t **audit_off);
/* hb_job_release of n jobs in one call (no counterpart in the reference). */
hb_status hb_jobs_release(hb_ctx *ctx, uint32_t n, const uint64_t *job_ids);

/* Synchronous convenience: submit + wait + copy results into malloc'd
 * arrays the caller frees with hb_free. */
hb_status hb_block(hb_ctx *ctx, const hb_plan *plan, const hb_table *l,
                   const hb_table *r, uint32_t **pairs, uint64_t *n_pairs,
                   hb_audit_entry **audit, uint64_t *n_audit, hb_job_stats *stats);
void hb_free(void *p);

/* One blocking job that allocates nothing.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  The job of hb_block on tables already prepared for the
 * plan (every signature, code and join layout it needs built by an earlier
 * regular job on them; else HB_ERR_INVALID_ARG), run synchronously with all
 * working memory taken from one static device array of the library (64 MiB
 * per GPU, held until the process ends) and no hipMalloc / hipFree.  The
 * pairs (2 * n_pairs uint32, (l, r)) and audit entries are written to the
 * caller's device buffers d_pairs / d_audit (pair_cap pairs, audit_cap
 * entries); *n_pairs / *n_audit receive the counts even when they exceed
 * the capacities, which returns HB_ERR_CAPACITY (so does a job whose pair
 * lists do not fit the static array).  Calls that overlap on one GPU share
 * the array and race: serialise them. */
hb_status hb_block_noalloc(hb_ctx *ctx, const hb_plan *plan, const hb_table *l,
                           const hb_table *r, uint32_t *d_pairs, uint64_t pair_cap,
                           hb_audit_entry *d_audit, uint64_t audit_cap, uint64_t *n_pairs,
                           uint64_t *n_audit);

/* ---- serving loop ----------------------------------------------------------
 * No counterpart in the reference.  A server owns a worker thread that
 * admits queued jobs into batches (at most max_jobs_per_batch jobs, at most
 * max_batches_in_flight batches running) while their estimated working
 * memory fits memory_budget bytes (0: three quarters of the device's free
 * memory at creation); a job that arrives while others run joins the next
 * batch and overlaps with them on the device.  Each job is a regular job of
 * the context: query, wait, read and release it by id (waiting covers the
 * time it is queued).  A job's result is the same alone and in any batch. */
typedef struct {
  int32_t max_batches_in_flight;
  int32_t max_jobs_per_batch;
  uint64_t memory_budget;
} hb_server_config;

hb_status hb_server_create(hb_ctx *ctx, const hb_server_config *cfg, hb_server **out);
hb_status hb_server_submit(hb_server *s, const hb_plan *plan, const hb_table *l,
                           const hb_table *r, uint64_t job_id);
/* Batches launched, jobs admitted so far, jobs waiting in the queue. */
hb_status hb_server_info(hb_server *s, uint64_t *batches, uint64_t *jobs, uint64_t *queued);
/* Runs the queue to the end, waits for running batches, stops the worker.
 * Jobs stay readable through the context until released. */
hb_status hb_server_destroy(hb_server *s);

/* Autotuner of the bound filter's launch shape: runs the given jobs once to
 * build their per-table structures, then twice per candidate shape, and
 * sets the shape with the smallest filter-stage time as the context's.
 * candidates holds n_candidates triples (rows per lane, tile_b, groups per
 * CU); NULL tries 24 built-in shapes.  Writes the chosen triple and its
 * time.  No counterpart in the reference (its launch is fixed). */
hb_status hb_autotune_filter_shape(hb_ctx *ctx, uint32_t n_jobs, const hb_plan *const *plans,
                                   const hb_table *const *l, const hb_table *const *r,
                                   const int32_t *candidates, uint32_t n_candidates,
                                   int32_t best[3], double *best_ms);

/* ---- host front end (no device needed) ---------------------------------- */

/* ExecutionPlanGenerator(rule_dir) + GetExecutionPlan(kEqualitiesFirst)
 * (execution_plan_generator.h:64-155): loads every regular file of rule_dir
 * (names sorted), decodes each as a rule (yaml_config.h:19-57), orders the
 * rules by number of equality entries, descending (stable), and serialises
 * equalities first.  The returned view stays valid until hb_host_plan_free. */
hb_status hb_rules_load(const char *rule_dir, hb_host_plan **out);
/* Same, from YAML documents given as strings (one rule per string). */
hb_status hb_rules_parse(int32_t n_docs, const char *const *docs, hb_host_plan **out);
hb_status hb_host_plan_view(const hb_host_plan *p, hb_serialized_plan *view);
void hb_host_plan_free(hb_host_plan *p);

/* rapidcsv::Document(path, LabelParams(-1, -1), SeparatorParams(sep))
 * (data_mngr.h:48-51): every line is a data row (no header), cells split on
 * sep outside double quotes, quotes removed ("" -> "), CR of CRLF dropped. */
hb_status hb_csv_load(const char *path, char sep, hb_host_table **out);
hb_status hb_host_table_view(const hb_host_table *t, hb_table_view *view);
/* Cell text (NUL-terminated, library-owned). */
const char *hb_host_table_cell(const hb_host_table *t, int64_t row, int32_t col);
void hb_host_table_free(hb_host_table *t);

/* DataPartitioning for two tables (data_mngr.h:81-222): bucket by
 * std::hash<std::string> of the column in plan slot 0 when that slot is an
 * equality (the same column index on both sides), optional first-character
 * sub-bucket (prefix_idx; pass HB_CHECK_POINT for none), partitions with an
 * empty side dropped.  Partition k holds its L and R row numbers in input
 * order. */
hb_status hb_partition(const hb_host_table *l, const hb_host_table *r,
                       const hb_serialized_plan *plan, uint64_t n_partitions,
                       int64_t prefix_idx, hb_partitions **out);
/* The same partitioning over column-addressed views (for tables that never
 * went through the CSV reader); identical buckets for identical cells. */
hb_status hb_partition_views(const hb_table_view *l, const hb_table_view *r,
                             const hb_serialized_plan *plan, uint64_t n_partitions,
                             int64_t prefix_idx, hb_partitions **out);
int64_t hb_partitions_count(const hb_partitions *p);
hb_status hb_partitions_get(const hb_partitions *p, int64_t k, const int64_t **l_rows,
                            int64_t *n_l, const int64_t **r_rows, int64_t *n_r);
void hb_partitions_free(hb_partitions *p);

/* WriteMatch (host_reducer.h:80-90): one line "<L[l][0]>,<R[r][0]>\n" per
 * pair, rows mapped through l_rows / r_rows when not NULL. */
hb_status hb_write_pairs(const char *path, int32_t append, const hb_host_table *l,
                         const hb_host_table *r, const int64_t *l_rows,
                         const int64_t *r_rows, const uint32_t *pairs, uint64_t n_pairs);

/* ---- filter coefficients (exposed for host-side verification) -----------
 * The bound filter rejects a candidate pair of cell lengths S and T (1..255)
 * with match upper bound m (sum over 32 byte classes of the smaller count)
 * when, in float arithmetic, (2m) * (1/S + 1/T) < 2 * coef[q], where q = 0
 * if the first bytes differ, q = 3 if a later one of the first 4 bytes
 * differs, and no pair with equal first 4 bytes is rejected.  coef[p] is the
 * Jaro bound 3 * J - 1 at which a common prefix of p (the largest allowed by
 * q) could still reach threshold - 2e-6, lowered by 1e-6 relative, so every
 * rejected pair has Jaro-Winkler below threshold - 2e-6. */
hb_status hb_filter_coefs(float threshold, float coef[4]);

/* ---- device arithmetic check ---------------------------------------------
 * Evaluates on the GPU, for n tuples mstp[4*i..4*i+3] = (m, S, T, p), the
 * Jaro tail of lev_jaro_ratio (device_func.cuh:127-128) followed by the
 * Winkler step of lev_jaro_winkler_ratio (device_func.cuh:229-230), exactly
 * as the blocking kernels do: out[i] unfused as written, out_fused[i] with
 * the last multiply-add fused.  Host arrays; synchronous. */
hb_status hb_jw_from_counts(hb_ctx *ctx, const int32_t *mstp, int64_t n, double *out,
                            double *out_fused);

#ifdef __cplusplus
}
#endif

#endif /* HYPERBLOCKER_AMD_H_ */
