/*
 * hb_oracle.c -- CPU restatement of HyperBlocker's blocking hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product; see hb_oracle.h).
 *
 * Build: gcc -O3 -ffp-contract=off (no -ffast-math), x86-64 SSE2 arithmetic,
 * so every float and double operation below rounds once, in source order.
 *
 * Pinning (what checks this file against something other than itself):
 *   - hand-derived known answers for the Jaro/JW restatement, chosen so that
 *     the textbook window, a transposition term, a Winkler prefix capped at 4,
 *     the dead host variant's "match + 1", or an all-double tail each fail one;
 *   - hbo_jaro_windowed, a second formulation of the greedy match written
 *     from the window definition rather than the two reference loops, compared
 *     exhaustively on short strings over small alphabets;
 *   - closed forms: identical strings -> 1, disjoint alphabets -> 0, empty
 *     cases, JW == 1 once the common prefix reaches 10, swap invariance of
 *     Jaro when lengths differ;
 *   - the plan loop against OR-of-ANDs over every truth assignment of small
 *     plans.
 * The reference itself cannot be built here (CUDA-only headers, empty
 * submodules), so there is no compiled reference to compare with.
 */
#define _GNU_SOURCE
#include "hb_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- per-pair predicates ------------------------------------------------ */

/* equalities_kernel, core/gpu/device_func.cuh:39-52: byte equality up to the
 * first NUL of either side. */
int hbo_equal(const char *a, const char *b) {
  int i = 0;
  while (a[i] && b[i]) {
    if (a[i] != b[i]) return 0;
    i++;
  }
  if (a[i]) return 0;
  if (b[i]) return 0;
  return 1;
}

/* Greedy match count of lev_jaro_ratio, device_func.cuh:71-111.
 * term_l is the inner (shorter, or left on a tie) string.  The reference's
 * index array has 256 slots and loop 2 can touch slot 256 when the inner
 * string has >= 257 bytes (device_func.cuh:83,103); that slot is modelled
 * here as a 257th zero-initialised slot. */
static size_t jaro_greedy(const char *term_l, size_t len_l, const char *term_r,
                          size_t len_r) {
  uint16_t idx[257];
  size_t i, j, halflen, match = 0, to;
  memset(idx, 0, sizeof(idx));
  halflen = (len_l + 1) / 2;
  /* loop 1, device_func.cuh:87-98 */
  for (i = 0; i < halflen; i++) {
    for (j = 0; j <= i + halflen; j++) {
      if (j >= 256) break;
      if (term_l[j] == term_r[i] && !idx[j]) {
        match++;
        idx[j] = (uint16_t)match;
        break;
      }
    }
  }
  /* loop 2, device_func.cuh:99-111 */
  to = len_l + halflen < len_r ? len_l + halflen : len_r;
  for (i = halflen; i < to; i++) {
    for (j = i - halflen; j < len_l; j++) {
      if (j > 256) break;
      if (term_l[j] == term_r[i] && !idx[j]) {
        match++;
        idx[j] = (uint16_t)match;
        break;
      }
    }
  }
  return match;
}

/* device_func.cuh:127-128: md = (float)m; (md/len_l + md/len_r + 1.0)/3.0,
 * i.e. two fp32 divides, an fp32 add, a double add of 1.0, a double divide
 * by 3.0, and the store into the fp32 result. */
static float jaro_tail(size_t match, size_t len_l, size_t len_r) {
  float md = (float)match;
  float q1 = md / (float)len_l;
  float q2 = md / (float)len_r;
  float s12 = q1 + q2;
  double d = (double)s12 + 1.0;
  d = d / 3.0;
  return (float)d;
}

/* lev_jaro_ratio, device_func.cuh:54-132. */
float hbo_jaro(const char *term_l, const char *term_r) {
  size_t len_l = strlen(term_l), len_r = strlen(term_r);
  if (len_r == 0 || len_l == 0) {
    if (len_l == 0 && len_r == 0) return 1.0f;
    return 0.0f;
  }
  if (len_l > len_r) { /* device_func.cuh:71-79: only a strict > swaps */
    const char *b = term_l;
    size_t t = len_l;
    term_l = term_r;
    term_r = b;
    len_l = len_r;
    len_r = t;
  }
  size_t match = jaro_greedy(term_l, len_l, term_r, len_r);
  if (!match) return 0.0f;
  return jaro_tail(match, len_l, len_r);
}

static size_t common_prefix(const char *a, size_t la, const char *b,
                            size_t lb) {
  size_t m = la < lb ? la : lb, p;
  for (p = 0; p < m; p++)
    if (a[p] != b[p]) break;
  return p;
}

/* lev_jaro_winkler_ratio, device_func.cuh:210-231:
 * j += (1.0 - j) * p * pfweight, evaluated left to right, then clamp at 1. */
static double winkler(float jaro, size_t p, int fused) {
  double j = (double)jaro;
  double x = (1.0 - j) * (double)p;
  if (fused)
    j = fma(x, 0.1, j);
  else
    j = j + x * 0.1;
  return j > 1.0 ? 1.0 : j;
}

double hbo_jw(const char *a, const char *b) {
  size_t la = strlen(a), lb = strlen(b);
  return winkler(hbo_jaro(a, b), common_prefix(a, la, b, lb), 0);
}

double hbo_jw_fused(const char *a, const char *b) {
  size_t la = strlen(a), lb = strlen(b);
  return winkler(hbo_jaro(a, b), common_prefix(a, la, b, lb), 1);
}

void hbo_jaro_counts(const char *a, const char *b, int32_t *m, int32_t *S,
                     int32_t *T, int32_t *p) {
  size_t la = strlen(a), lb = strlen(b);
  *p = (int32_t)common_prefix(a, la, b, lb);
  if (la == 0 || lb == 0) {
    *m = 0;
    *S = (int32_t)(la < lb ? la : lb);
    *T = (int32_t)(la < lb ? lb : la);
    return;
  }
  if (la > lb)
    *m = (int32_t)jaro_greedy(b, lb, a, la);
  else
    *m = (int32_t)jaro_greedy(a, la, b, lb);
  *S = (int32_t)(la < lb ? la : lb);
  *T = (int32_t)(la < lb ? lb : la);
}

double hbo_jw_from_counts(int32_t m, int32_t S, int32_t T, int32_t p) {
  float jaro;
  if (S == 0 || T == 0)
    jaro = (S == 0 && T == 0) ? 1.0f : 0.0f;
  else if (m == 0)
    jaro = 0.0f;
  else
    jaro = jaro_tail((size_t)m, (size_t)S, (size_t)T);
  return winkler(jaro, (size_t)p, 0);
}

void hbo_jw_from_counts_batch(const int32_t *mstp, int64_t n, double *out) {
  for (int64_t i = 0; i < n; i++)
    out[i] = hbo_jw_from_counts(mstp[4 * i], mstp[4 * i + 1], mstp[4 * i + 2], mstp[4 * i + 3]);
}

/* Second formulation: for every outer position i < min(T, S + h), take the
 * lowest free inner position j with s[j] == t[i] inside the window
 * [max(0, i-h), min(i+h, S-1)], where for S >= 257 loop 1 (i < h) stops
 * below 256 and loop 2 stops at 256. */
float hbo_jaro_windowed(const char *a, const char *b) {
  size_t la = strlen(a), lb = strlen(b);
  if (la == 0 || lb == 0) return (la == 0 && lb == 0) ? 1.0f : 0.0f;
  const char *s = a, *t = b;
  size_t S = la, T = lb;
  if (la > lb) {
    s = b;
    t = a;
    S = lb;
    T = la;
  }
  long h = (long)(S + 1) / 2;
  long n_out = (long)T < (long)S + h ? (long)T : (long)S + h;
  uint64_t freebits[5] = {0, 0, 0, 0, 0}; /* 320 >= 257 positions */
  for (long j = 0; j < (long)S && j < 320; j++)
    freebits[j >> 6] |= 1ull << (j & 63);
  long m = 0;
  for (long i = 0; i < n_out; i++) {
    long lo = i - h > 0 ? i - h : 0;
    long hi = i + h < (long)S - 1 ? i + h : (long)S - 1;
    long cap = i < h ? 255 : 256;
    if (hi > cap) hi = cap;
    for (long j = lo; j <= hi; j++) {
      if ((freebits[j >> 6] >> (j & 63)) & 1ull) {
        if (s[j] == t[i]) {
          freebits[j >> 6] &= ~(1ull << (j & 63));
          m++;
          break;
        }
      }
    }
  }
  if (m == 0) return 0.0f;
  return jaro_tail((size_t)m, S, T);
}

uint64_t hbo_jw_work(const char *a, const char *b) {
  size_t la = strlen(a), lb = strlen(b);
  size_t S = la < lb ? la : lb, T = la < lb ? lb : la;
  size_t h = (S + 1) / 2;
  size_t n_out = T < S + h ? T : S + h;
  size_t p = common_prefix(a, la, b, lb);
  return 2 * (uint64_t)S + 8 * (uint64_t)n_out * ((S + 31) / 32) + (p + 1) +
         40;
}

/* ---- plan --------------------------------------------------------------- */

int hbo_plan_check(const hbo_plan *plan) {
  if (plan->length <= 0 || plan->n_rules <= 0) return -1;
  if (plan->pred_index[plan->length - 1] != HBO_CHECK_POINT) return -2;
  int i = 0, rules = 0;
  while (i < plan->length) {
    if (plan->pred_index[i] == HBO_CHECK_POINT) return -3; /* empty rule */
    while (i < plan->length && plan->pred_index[i] != HBO_CHECK_POINT) {
      if (i + 1 >= plan->length) return -2;
      char t0 = plan->pred_type[i], t1 = plan->pred_type[i + 1];
      if (t0 != 'e' && t0 != 's') return -4;
      if (t0 != t1) return -5;
      if (plan->pred_index[i] < 0 || plan->pred_index[i + 1] < 0) return -6;
      if (plan->pred_index[i + 1] == HBO_CHECK_POINT) return -2;
      i += 2;
    }
    if (i >= plan->length) return -2;
    i++; /* checkpoint */
    rules++;
  }
  return rules == plan->n_rules ? 0 : -1;
}

static inline const char *cell(const hbo_table *t, int64_t r, int32_t c) {
  return t->col_base[c] + r * t->col_stride[c];
}

/* global_func.cuh:118-167, one pair (row_index_l, row_index_r). */
int hbo_eval_pair(const hbo_plan *plan, const hbo_table *L, int64_t l,
                  const hbo_table *R, int64_t r, hbo_counters *cnt) {
  const int32_t *pi = plan->pred_index;
  const int length = plan->length;
  int is_match = 0;
  if (cnt) cnt->pairs++;
  for (int i = 0; i < length; i += 2) {
    int to_skip = 0;
    int il = pi[i], ir = pi[i + 1];
    switch (plan->pred_type[i]) {
    case 'e':
      if (cnt) {
        cnt->eq_evals++;
        cnt->work += 1;
      }
      if (!hbo_equal(cell(L, l, il), cell(R, r, ir))) to_skip = 1;
      break;
    case 's': {
      const char *a = cell(L, l, il), *b = cell(R, r, ir);
      if (cnt) {
        cnt->jw_evals++;
        cnt->work += hbo_jw_work(a, b);
      }
      double sim = hbo_jw(a, b);
      if (sim < (double)plan->pred_threshold[i]) to_skip = 1;
      break;
    }
    default:
      break;
    }
    if (to_skip) { /* move to the slot after the next checkpoint */
      while (i + 2 < length) {
        if (pi[i + 2] == HBO_CHECK_POINT) {
          i++;
          break;
        }
        i += 2;
      }
    } else if (i + 2 < length && pi[i + 2] == HBO_CHECK_POINT) {
      is_match = 1;
      break;
    }
  }
  if (is_match && cnt) cnt->matches++;
  return is_match;
}

/* ---- threaded drivers --------------------------------------------------- */

typedef struct {
  void *data;
  int64_t n, cap;
} vec_t;

static int vec_push(vec_t *v, const void *item, size_t sz) {
  if (v->n == v->cap) {
    int64_t nc = v->cap ? v->cap * 2 : 256;
    void *nd = realloc(v->data, (size_t)nc * sz);
    if (!nd) return -1;
    v->data = nd;
    v->cap = nc;
  }
  memcpy((char *)v->data + (size_t)v->n * sz, item, sz);
  v->n++;
  return 0;
}

#define CHUNK_ROWS 64

typedef struct {
  const hbo_plan *plan;
  const hbo_table *L, *R;
  int64_t l_begin, l_step, n_sampled;
  int64_t n_chunks;
  int64_t next_chunk; /* atomic */
  vec_t *chunk_out;   /* one vector per chunk */
  int mode;           /* 0 = pairs, 1 = audit */
  int failed;
  pthread_mutex_t mu;
  hbo_counters total;
} job_t;

/* Each rule on its own, short-circuit inside the rule; record similarity
 * slots whose JW is within 1e-6 of the threshold. */
static int audit_pair(const hbo_plan *plan, const hbo_table *L, int64_t l,
                      const hbo_table *R, int64_t r, vec_t *out) {
  const int32_t *pi = plan->pred_index;
  int i = 0;
  while (i < plan->length) {
    int alive = 1;
    while (i < plan->length && pi[i] != HBO_CHECK_POINT) {
      if (alive) {
        const char *a = cell(L, l, pi[i]), *b = cell(R, r, pi[i + 1]);
        if (plan->pred_type[i] == 'e') {
          alive = hbo_equal(a, b);
        } else {
          double thr = (double)plan->pred_threshold[i];
          double jw = hbo_jw(a, b);
          if (fabs(jw - thr) < 1e-6) {
            hbo_audit_entry e;
            double jf = hbo_jw_fused(a, b);
            size_t la = strlen(a), lb = strlen(b);
            e.l = (uint32_t)l;
            e.r = (uint32_t)r;
            e.slot = i;
            e.flags = ((jw < thr) != (jf < thr)) ? 1 : 0;
            if ((la < lb ? la : lb) >= 257) e.flags |= 2;
            e.jw = jw;
            if (vec_push(out, &e, sizeof(e))) return -1;
          }
          alive = !(jw < thr);
        }
      }
      i += 2;
    }
    i++;
  }
  return 0;
}

static void *worker(void *arg) {
  job_t *job = (job_t *)arg;
  hbo_counters local;
  memset(&local, 0, sizeof(local));
  for (;;) {
    int64_t c = __atomic_fetch_add(&job->next_chunk, 1, __ATOMIC_RELAXED);
    if (c >= job->n_chunks) break;
    vec_t *v = &job->chunk_out[c];
    int64_t s0 = c * CHUNK_ROWS;
    int64_t s1 = s0 + CHUNK_ROWS < job->n_sampled ? s0 + CHUNK_ROWS
                                                   : job->n_sampled;
    for (int64_t s = s0; s < s1; s++) {
      int64_t l = job->l_begin + s * job->l_step;
      for (int64_t r = 0; r < job->R->n_rows; r++) {
        if (job->mode == 0) {
          if (hbo_eval_pair(job->plan, job->L, l, job->R, r, &local)) {
            uint32_t pr[2] = {(uint32_t)l, (uint32_t)r};
            if (vec_push(v, pr, sizeof(pr))) job->failed = 1;
          }
        } else {
          if (audit_pair(job->plan, job->L, l, job->R, r, v)) job->failed = 1;
        }
      }
    }
  }
  pthread_mutex_lock(&job->mu);
  job->total.pairs += local.pairs;
  job->total.eq_evals += local.eq_evals;
  job->total.jw_evals += local.jw_evals;
  job->total.work += local.work;
  job->total.matches += local.matches;
  pthread_mutex_unlock(&job->mu);
  return NULL;
}

static int64_t run_job(const hbo_plan *plan, const hbo_table *L,
                       const hbo_table *R, int64_t l_begin, int64_t l_step,
                       int n_threads, int mode, void **out, size_t item_sz,
                       hbo_counters *cnt) {
  *out = NULL;
  if (hbo_plan_check(plan)) return -1;
  if (l_step < 1 || l_begin < 0) return -2;
  for (int i = 0; i < plan->length;) { /* slot pairs index (L, R) columns */
    if (plan->pred_index[i] == HBO_CHECK_POINT) {
      i++;
      continue;
    }
    if (plan->pred_index[i] >= L->n_cols || plan->pred_index[i + 1] >= R->n_cols)
      return -4;
    i += 2;
  }
  job_t job;
  memset(&job, 0, sizeof(job));
  job.plan = plan;
  job.L = L;
  job.R = R;
  job.l_begin = l_begin;
  job.l_step = l_step;
  job.n_sampled = l_begin < L->n_rows ? (L->n_rows - l_begin + l_step - 1) / l_step : 0;
  job.n_chunks = (job.n_sampled + CHUNK_ROWS - 1) / CHUNK_ROWS;
  job.mode = mode;
  pthread_mutex_init(&job.mu, NULL);
  job.chunk_out = (vec_t *)calloc((size_t)(job.n_chunks ? job.n_chunks : 1), sizeof(vec_t));
  if (!job.chunk_out) return -3;
  if (n_threads < 1) n_threads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
  for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, worker, &job);
  for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
  free(th);
  int64_t total = 0;
  for (int64_t c = 0; c < job.n_chunks; c++) total += job.chunk_out[c].n;
  char *res = (char *)malloc((size_t)(total ? total : 1) * item_sz);
  int64_t pos = 0;
  for (int64_t c = 0; c < job.n_chunks; c++) {
    if (res && job.chunk_out[c].n)
      memcpy(res + (size_t)pos * item_sz, job.chunk_out[c].data,
             (size_t)job.chunk_out[c].n * item_sz);
    pos += job.chunk_out[c].n;
    free(job.chunk_out[c].data);
  }
  free(job.chunk_out);
  pthread_mutex_destroy(&job.mu);
  if (cnt) *cnt = job.total;
  if (!res || job.failed) {
    free(res);
    return -3;
  }
  *out = res;
  return total;
}

int64_t hbo_block(const hbo_plan *plan, const hbo_table *L, const hbo_table *R,
                  int64_t l_begin, int64_t l_step, int n_threads,
                  uint32_t **out, hbo_counters *cnt) {
  /* Rows inside a chunk are visited in increasing l, then r, and chunks are
   * concatenated in order, so the result is sorted by (l, r). */
  return run_job(plan, L, R, l_begin, l_step, n_threads, 0, (void **)out,
                 2 * sizeof(uint32_t), cnt);
}

int64_t hbo_audit(const hbo_plan *plan, const hbo_table *L,
                  const hbo_table *R, int64_t l_begin, int64_t l_step,
                  int n_threads, hbo_audit_entry **out) {
  return run_job(plan, L, R, l_begin, l_step, n_threads, 1, (void **)out,
                 sizeof(hbo_audit_entry), NULL);
}

/* hbo_eval_pair over a list of pairs, on n_threads threads. */
typedef struct {
  const hbo_plan *plan;
  const hbo_table *L, *R;
  const uint32_t *pairs;
  int64_t n, next; /* next: atomic chunk cursor */
  uint8_t *out;
} eval_job_t;

static void *eval_worker(void *arg) {
  eval_job_t *job = (eval_job_t *)arg;
  for (;;) {
    int64_t b = __atomic_fetch_add(&job->next, 4096, __ATOMIC_RELAXED);
    if (b >= job->n) break;
    int64_t e = b + 4096 < job->n ? b + 4096 : job->n;
    for (int64_t i = b; i < e; i++)
      job->out[i] = (uint8_t)hbo_eval_pair(job->plan, job->L, job->pairs[2 * i], job->R,
                                           job->pairs[2 * i + 1], NULL);
  }
  return NULL;
}

int hbo_eval_pairs(const hbo_plan *plan, const hbo_table *L, const hbo_table *R,
                   const uint32_t *pairs, int64_t n, int n_threads, uint8_t *out) {
  if (hbo_plan_check(plan)) return -1;
  for (int64_t i = 0; i < n; i++)
    if ((int64_t)pairs[2 * i] >= L->n_rows || (int64_t)pairs[2 * i + 1] >= R->n_rows) return -2;
  eval_job_t job = {plan, L, R, pairs, n, 0, out};
  if (n_threads < 1) n_threads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
  if (!th) return -3;
  for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, eval_worker, &job);
  for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
  free(th);
  return 0;
}

void hbo_free(void *p) { free(p); }
