/*
 * hb_oracle.h -- CPU restatement of HyperBlocker's rule-based blocking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in hyperblocker_amd/
 * links, imports or calls it.
 *
 * Every function cites the reference file:line it restates.  The reference
 * source is read as text; none of it is copied.
 */
#ifndef HB_ORACLE_H_
#define HB_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column-addressed table view.  Cell (r, c) is the NUL-terminated byte string
 * at col_base[c] + r * col_stride[c].  The reference's row-major
 * SerializedTable (core/data_structures/table.h:16-51) is the special case
 * col_base[c] = data + col_offset[c], col_stride[c] = aligned_tuple_size. */
typedef struct {
  int64_t n_rows;
  int32_t n_cols;
  const char *const *col_base;
  const int64_t *col_stride;
} hbo_table;

/* Same fields and meaning as SerializedExecutionPlan
 * (core/components/execution_plan_generator.h:25-30). */
typedef struct {
  int32_t n_rules;
  int32_t length;
  const int32_t *pred_index;
  const char *pred_type;
  const float *pred_threshold;
} hbo_plan;

typedef struct {
  uint64_t pairs;    /* pairs whose plan was interpreted */
  uint64_t eq_evals; /* equalities_kernel calls, short-circuit order */
  uint64_t jw_evals; /* lev_jaro_winkler_ratio calls, short-circuit order */
  uint64_t work;     /* cost-model int32 lane-ops (see hbo_jw_work) */
  uint64_t matches;
} hbo_counters;

typedef struct {
  uint32_t l, r;
  int32_t slot;  /* plan slot of the similarity predicate */
  int32_t flags; /* bit0: fused multiply-add tail decides differently;
                    bit1: shorter string has >= 257 bytes */
  double jw;
} hbo_audit_entry;

#define HBO_CHECK_POINT 2147483647

/* equalities_kernel, core/gpu/device_func.cuh:39-52 */
int hbo_equal(const char *a, const char *b);

/* lev_jaro_ratio, core/gpu/device_func.cuh:54-132 (loop-exact). */
float hbo_jaro(const char *a, const char *b);

/* lev_jaro_winkler_ratio, core/gpu/device_func.cuh:210-231, pfweight 0.1,
 * evaluated unfused as written. */
double hbo_jw(const char *a, const char *b);

/* Same as hbo_jw but with the last multiply-add of device_func.cuh:229
 * fused into one rounding (what nvcc's default --fmad=true may emit). */
double hbo_jw_fused(const char *a, const char *b);

/* Counts behind hbo_jaro: greedy matches m, S = shorter length, T = longer
 * length, p = common prefix of (a, b) in argument order. */
void hbo_jaro_counts(const char *a, const char *b, int32_t *m, int32_t *S,
                     int32_t *T, int32_t *p);

/* The float/double tail of lev_jaro_ratio + lev_jaro_winkler_ratio as a
 * function of the counts (device_func.cuh:113-128, 222-230). */
double hbo_jw_from_counts(int32_t m, int32_t S, int32_t T, int32_t p);

/* hbo_jw_from_counts over n tuples mstp[4*i..4*i+3] = (m, S, T, p). */
void hbo_jw_from_counts_batch(const int32_t *mstp, int64_t n, double *out);

/* An independently structured formulation of the same greedy matching: a
 * symmetric window of radius h around each outer position and a bitmask of
 * free inner positions.  Used only to cross-check hbo_jaro. */
float hbo_jaro_windowed(const char *a, const char *b);

/* Cost-model work of one JW evaluation (int32 lane-ops):
 * 2*S + 8*n_out*ceil(S/32) + (p+1) + 40, n_out = min(T, S + (S+1)/2). */
uint64_t hbo_jw_work(const char *a, const char *b);

/* 0 if the plan is well formed, else a negative code: -1 empty/negative
 * length, -2 missing trailing checkpoint, -3 empty rule, -4 bad predicate
 * type, -5 slot pair with mismatched types, -6 negative column index. */
int hbo_plan_check(const hbo_plan *plan);

/* blocking_kernel's per-pair plan loop, core/gpu/global_func.cuh:118-167.
 * Returns 1 if the pair is emitted. */
int hbo_eval_pair(const hbo_plan *plan, const hbo_table *L, int64_t l,
                  const hbo_table *R, int64_t r, hbo_counters *cnt);

/* blocking_kernel over one partition, global_func.cuh:104-185, with the
 * 65,536 output cap removed.  Left rows l_begin, l_begin + l_step, ... are
 * evaluated against every right row.  n_threads worker threads take chunks
 * of 64 left rows.  *out receives malloc'd (l, r) uint32 pairs sorted by
 * (l, r); free with hbo_free.  Returns the pair count or a negative error. */
int64_t hbo_block(const hbo_plan *plan, const hbo_table *L, const hbo_table *R,
                  int64_t l_begin, int64_t l_step, int n_threads,
                  uint32_t **out, hbo_counters *cnt);

/* Audit list: every (pair, similarity slot) whose JW lies within 1e-6 of the
 * slot's threshold, where each rule is evaluated on its own with
 * short-circuiting inside the rule.  Sorted by (l, r, slot). */
int64_t hbo_audit(const hbo_plan *plan, const hbo_table *L,
                  const hbo_table *R, int64_t l_begin, int64_t l_step,
                  int n_threads, hbo_audit_entry **out);

/* hbo_eval_pair for each of n (l, r) pairs (2*n uint32), on n_threads
 * threads: out[i] = 1 if pair i is emitted.  Returns 0, or a negative error
 * (malformed plan, row out of range). */
int hbo_eval_pairs(const hbo_plan *plan, const hbo_table *L, const hbo_table *R,
                   const uint32_t *pairs, int64_t n, int n_threads, uint8_t *out);

void hbo_free(void *p);

#ifdef __cplusplus
}
#endif

#endif /* HB_ORACLE_H_ */
