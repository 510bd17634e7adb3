// gfx950 kernels of the rule-based blocking path.
//
// Reference semantics: core/gpu/global_func.cuh:104-185 (blocking_kernel),
// core/gpu/device_func.cuh:39-52 (equalities_kernel), :54-132
// (lev_jaro_ratio), :210-231 (lev_jaro_winkler_ratio).
//
// Structure of one rule's evaluation on the device:
//   1. join: the first equality's dictionary codes of both sides are sorted
//      (hb_sort.hip); each run of one code on the left meets the run of the
//      same code on the right (k_run_info).  A rule without equalities is one
//      group covering L x R.
//   2. k_filter_groups: pairs of every group are checked for the remaining
//      equalities (integer code compares) and an upper bound of the first
//      similarity predicate (byte-class histograms, v_sad_u8); survivors are
//      compacted with wave ballots into a pair list.
//   3. k_exact: each surviving pair gets the exact greedy Jaro match count
//      (one wave per pair: a v_cmp per outer character gives the equal
//      positions of the inner string as a 64-bit mask, the greedy pick runs on
//      the scalar unit) and the reference's fp32/fp64 tail; later similarity
//      predicates of the rule run as further k_exact stages on the list.
//
// Floating point: contraction is off for this file, fp32 '/' is correctly
// rounded (hipcc default), so every operation of the Jaro/Winkler tail
// rounds once, in the reference's order.
#include <hip/hip_runtime.h>

#include "../../include/hyperblocker_amd.h"
#include "hb_internal.h"

#pragma clang fp contract(off)

namespace hb {
namespace {

constexpr unsigned long long kEmpty = ~0ull;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}

// Appends (l, r) of every lane with ok set; one atomic per wave.
__device__ __forceinline__ void wave_append(bool ok, uint32_t l, uint32_t r, const StageOut &out) {
  const uint64_t mask = __ballot(ok);
  if (mask == 0) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(out.count, (unsigned long long)__popcll(mask));
  base = readlane64(base, leader);
  if (ok) {
    const unsigned long long pos = base + rank_in(mask);
    if (pos < out.cap) {
      out.pairs[2 * pos] = l;
      out.pairs[2 * pos + 1] = r;
    } else {
      *out.overflow = 1ull;
    }
  }
}

// lev_jaro_ratio tail, device_func.cuh:127-128.
__device__ __forceinline__ float jaro_tail(int m, int S, int T) {
  const float md = (float)m;
  const float q1 = md / (float)S;
  const float q2 = md / (float)T;
  const float s12 = q1 + q2;
  double d = (double)s12 + 1.0;
  d = d / 3.0;
  return (float)d;
}

__device__ __forceinline__ float jaro_value(int m, int S, int T) {
  if (S == 0) return T == 0 ? 1.0f : 0.0f;
  if (m == 0) return 0.0f;
  return jaro_tail(m, S, T);
}

// lev_jaro_winkler_ratio, device_func.cuh:222-230.
__device__ __forceinline__ double winkler(float jaro, int p) {
  double j = (double)jaro;
  const double x = (1.0 - j) * (double)p;
  j = j + x * 0.1;
  return j > 1.0 ? 1.0 : j;
}

__device__ __forceinline__ double winkler_fused(float jaro, int p) {
  double j = (double)jaro;
  const double x = (1.0 - j) * (double)p;
  j = __builtin_fma(x, 0.1, j);
  return j > 1.0 ? 1.0 : j;
}

// Upper bound test of one similarity predicate.  Returns false only if
// the pair's Jaro-Winkler is certainly below threshold - 1e-6 (the filter
// coefficients carry a 2e-6 margin; see hb_filter_coefs).
__device__ __forceinline__ bool may_pass(const uint32_t (&ah)[8], uint32_t apre, uint32_t alen,
                                         const uint32_t (&bh)[8], uint32_t bpre, uint32_t blen,
                                         float c0, float c1, float c2, float c3) {
  const uint32_t la = alen & 0xffffu, lb = blen & 0xffffu;
  const uint32_t S = min(la, lb), T = max(la, lb);
  const uint32_t x = apre ^ bpre;
  uint32_t sad = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) sad = __builtin_amdgcn_sad_u8(ah[k], bh[k], sad);
  const uint32_t mub = (la + lb - sad) >> 1;  // sum over classes of min(count_a, count_b)
  const uint32_t p = ((uint32_t)(__ffs(x) - 1)) >> 3;
  const float coef = p == 0 ? c0 : (p == 1 ? c1 : (p == 2 ? c2 : c3));
  const bool reject = (float)(mub * (S + T)) < coef * (float)(S * T);
  return ((alen | blen) & kLongFlag) || S == 0 || x == 0 || !reject;
}

// ---- table preparation ---------------------------------------------------

__global__ void k_build_sig(const uint8_t *__restrict__ bytes, uint32_t width, int64_t n,
                            uint16_t *__restrict__ len_out, SigRow *__restrict__ sig_out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint4 *cell = reinterpret_cast<const uint4 *>(bytes + r * (int64_t)width);
  uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t L = 0, pre = 0;
  const uint32_t n16 = width >> 4;
  for (uint32_t k = 0; k < n16; k++) {
    const uint4 v = cell[k];
    if (k == 0) pre = v.x;
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    bool has_zero = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const uint32_t c = (w[q] >> (8 * b)) & 0xffu;
        has_zero |= (c == 0);
        const uint32_t cls = c & 31u;
        const uint32_t inc = (c != 0) ? (1u << ((cls & 3u) * 8u)) : 0u;
        L += (c != 0);
#pragma unroll
        for (int j = 0; j < 8; j++) h[j] += ((cls >> 2) == (uint32_t)j) ? inc : 0u;
      }
    }
    if (has_zero) break;
  }
  len_out[r] = (uint16_t)min(L, 65535u);
  SigRow s;
#pragma unroll
  for (int j = 0; j < 8; j++) s.h[j] = h[j];
  s.pre = pre;
  s.len = L | (L > 255u ? kLongFlag : 0u);
  s.pad[0] = s.pad[1] = 0;
  sig_out[r] = s;
}

__global__ void k_iota(uint32_t *out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long n, unsigned long long v) {
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_run_flags(const uint32_t *__restrict__ keys, uint32_t n, uint32_t *flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t *a, uint32_t lo, uint32_t hi,
                                                    uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] <= key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t *a, uint32_t lo, uint32_t hi,
                                                    uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void set_group_work(GroupArrays g, uint32_t gi, uint32_t l_cnt,
                                               uint32_t r_cnt) {
  const uint64_t pairs = (uint64_t)l_cnt * r_cnt;
  uint64_t small = 0, items = 0;
  if (pairs > kSmallGroupPairs) {
    const uint32_t a = max(l_cnt, r_cnt), b = min(l_cnt, r_cnt);
    items = (uint64_t)((a + kTileA - 1) / kTileA) * ((b + kTileB - 1) / kTileB);
  } else {
    small = pairs;
  }
  g.small_pairs[gi] = small;
  g.items[gi] = items;
}

__global__ void k_run_info(const uint32_t *__restrict__ sl, const uint32_t *__restrict__ flags,
                           const uint32_t *__restrict__ run_id, uint32_t n_l,
                           const uint32_t *__restrict__ sr, uint32_t n_r, GroupArrays g) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n_l) return;
  const uint32_t n_runs = n_l ? run_id[n_l - 1] + flags[n_l - 1] : 0u;
  if (i == 0) *g.n_runs = n_runs;
  if (i >= n_runs) {
    g.l_off[i] = g.l_cnt[i] = g.r_off[i] = g.r_cnt[i] = 0;
    g.small_pairs[i] = 0;
    g.items[i] = 0;
  }
  if (i < n_l && flags[i]) {
    const uint32_t gi = run_id[i];
    const uint32_t key = sl[i];
    const uint32_t l_end = upper_bound_u32(sl, i, n_l, key);
    const uint32_t r_lo = lower_bound_u32(sr, 0, n_r, key);
    const uint32_t r_hi = upper_bound_u32(sr, r_lo, n_r, key);
    g.l_off[gi] = i;
    g.l_cnt[gi] = l_end - i;
    g.r_off[gi] = r_lo;
    g.r_cnt[gi] = r_hi - r_lo;
    set_group_work(g, gi, l_end - i, r_hi - r_lo);
    const unsigned long long jp = (unsigned long long)(l_end - i) * (r_hi - r_lo);
    if (jp) atomicAdd(g.joined, jp);
  }
}

__global__ void k_single_group(uint32_t n_l, uint32_t n_r, GroupArrays g) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *g.n_runs = 1;
    g.l_off[0] = 0;
    g.l_cnt[0] = n_l;
    g.r_off[0] = 0;
    g.r_cnt[0] = n_r;
    set_group_work(g, 0, n_l, n_r);
    atomicAdd(g.joined, (unsigned long long)n_l * n_r);
    g.l_off[1] = g.l_cnt[1] = g.r_off[1] = g.r_cnt[1] = 0;
    g.small_pairs[1] = 0;
    g.items[1] = 0;
  }
}

// Largest g in [0, n) with off[g] <= x (off non-decreasing, off[0] = 0).
__device__ __forceinline__ uint32_t find_group(const uint64_t *off, uint32_t n, uint64_t x) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (off[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// ---- stage 1: equalities + bound filter over join groups ------------------

struct FilterArgs {
  GroupArrays g;
  const uint32_t *perm_l, *perm_r;
  const SigRow *sig_l, *sig_r;
  EqCheck eq;
  int32_t has_sim;
  float c0, c1, c2, c3;
  StageOut out;
};

// Work item = kTileA rows of the larger side in lanes (A) x up to kTileB
// rows of the other side (B), streamed through LDS in chunks of kChunkB.
__global__ __launch_bounds__(256) void k_filter_tiles(FilterArgs a) {
  __shared__ SigRow s_sig[kChunkB];
  __shared__ uint32_t s_row[kChunkB];
  __shared__ uint32_t s_code[kMaxEq][kChunkB];
  const uint32_t tid = threadIdx.x;
  const uint32_t n_runs = *a.g.n_runs;
  const uint64_t total = a.g.item_off[a.g.cap];
  for (uint64_t item = blockIdx.x; item < total; item += gridDim.x) {
    const uint32_t gi = find_group(a.g.item_off, n_runs, item);
    const uint32_t l_off = a.g.l_off[gi], l_cnt = a.g.l_cnt[gi];
    const uint32_t r_off = a.g.r_off[gi], r_cnt = a.g.r_cnt[gi];
    const bool a_is_l = l_cnt >= r_cnt;
    const uint32_t a_cnt = a_is_l ? l_cnt : r_cnt, b_cnt = a_is_l ? r_cnt : l_cnt;
    const uint32_t a_off = a_is_l ? l_off : r_off, b_off = a_is_l ? r_off : l_off;
    const uint32_t *perm_a = a_is_l ? a.perm_l : a.perm_r;
    const uint32_t *perm_b = a_is_l ? a.perm_r : a.perm_l;
    const SigRow *sig_a = a_is_l ? a.sig_l : a.sig_r;
    const SigRow *sig_b = a_is_l ? a.sig_r : a.sig_l;
    const uint32_t n_bseg = (b_cnt + kTileB - 1) / kTileB;
    const uint64_t local = item - a.g.item_off[gi];
    const uint32_t a_tile = (uint32_t)(local / n_bseg), b_seg = (uint32_t)(local % n_bseg);
    const uint32_t a_idx = a_tile * kTileA + tid;
    const bool a_valid = a_idx < a_cnt;
    const uint32_t a_row = a_valid ? perm_a[a_off + a_idx] : 0u;
    uint32_t ah[8] = {0, 0, 0, 0, 0, 0, 0, 0}, apre = 0, alen = 0;
    if (a_valid && a.has_sim) {
      const uint4 *src = reinterpret_cast<const uint4 *>(sig_a + a_row);
      const uint4 v0 = src[0], v1 = src[1], v2 = src[2];
      ah[0] = v0.x; ah[1] = v0.y; ah[2] = v0.z; ah[3] = v0.w;
      ah[4] = v1.x; ah[5] = v1.y; ah[6] = v1.z; ah[7] = v1.w;
      apre = v2.x;
      alen = v2.y;
    }
    uint32_t acode[kMaxEq];
#pragma unroll
    for (int k = 0; k < kMaxEq; k++)
      acode[k] = (k < a.eq.n && a_valid) ? (a_is_l ? a.eq.code_l[k] : a.eq.code_r[k])[a_row] : 0u;
    const uint32_t b_begin = b_seg * kTileB;
    const uint32_t b_end = min(b_begin + kTileB, b_cnt);
    for (uint32_t cb = b_begin; cb < b_end; cb += kChunkB) {
      __syncthreads();
      const uint32_t j = cb + tid;
      if (j < b_end) {
        const uint32_t b_row = perm_b[b_off + j];
        s_row[tid] = b_row;
        if (a.has_sim) s_sig[tid] = sig_b[b_row];
#pragma unroll
        for (int k = 0; k < kMaxEq; k++)
          if (k < a.eq.n) s_code[k][tid] = (a_is_l ? a.eq.code_r[k] : a.eq.code_l[k])[b_row];
      }
      __syncthreads();
      const uint32_t n_b = min((uint32_t)kChunkB, b_end - cb);
      for (uint32_t jj = 0; jj < n_b; jj++) {
        bool ok = a_valid;
#pragma unroll
        for (int k = 0; k < kMaxEq; k++)
          if (k < a.eq.n) ok = ok && (acode[k] == s_code[k][jj]);
        if (a.has_sim) {
          const SigRow &b = s_sig[jj];
          const uint32_t bh[8] = {b.h[0], b.h[1], b.h[2], b.h[3], b.h[4], b.h[5], b.h[6], b.h[7]};
          ok = ok && may_pass(ah, apre, alen, bh, b.pre, b.len, a.c0, a.c1, a.c2, a.c3);
        }
        const uint32_t b_row = s_row[jj];
        wave_append(ok, a_is_l ? a_row : b_row, a_is_l ? b_row : a_row, a.out);
      }
    }
  }
}

// Pairs of small groups, one pair per lane.
__global__ __launch_bounds__(256) void k_filter_small(FilterArgs a) {
  const uint32_t n_runs = *a.g.n_runs;
  const uint64_t total = a.g.pair_off[a.g.cap];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // Trip count is uniform per wave: every lane runs the same number of
  // rounds so the ballots inside see all lanes.
  const uint64_t base0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
  for (uint64_t q0 = base0; q0 < total; q0 += stride) {
    const uint64_t q = q0 + lane_id();
    bool ok = q < total;
    uint32_t l_row = 0, r_row = 0;
    if (ok) {
      const uint32_t gi = find_group(a.g.pair_off, n_runs, q);
      const uint64_t local = q - a.g.pair_off[gi];
      const uint32_t r_cnt = a.g.r_cnt[gi];
      const uint32_t li = (uint32_t)(local / r_cnt), ri = (uint32_t)(local % r_cnt);
      l_row = a.perm_l[a.g.l_off[gi] + li];
      r_row = a.perm_r[a.g.r_off[gi] + ri];
#pragma unroll
      for (int k = 0; k < kMaxEq; k++)
        if (k < a.eq.n) ok = ok && (a.eq.code_l[k][l_row] == a.eq.code_r[k][r_row]);
    }
    if (ok && a.has_sim) {
      const SigRow sa = a.sig_l[l_row], sb = a.sig_r[r_row];
      ok = may_pass(sa.h, sa.pre, sa.len, sb.h, sb.pre, sb.len, a.c0, a.c1, a.c2, a.c3);
    }
    wave_append(ok, l_row, r_row, a.out);
  }
}

// ---- exact Jaro-Winkler -----------------------------------------------------

// Greedy match count of lev_jaro_ratio (device_func.cuh:81-111) for one
// pair, computed by one wave.  Lane k holds inner bytes k + 64 w; for outer
// position i the window is [max(0, i-h), min(i+h, S-1)] (loop 1 stops below
// 256, loop 2 at 256), and the lowest free inner position with an equal byte
// is taken.  All control values are wave-uniform.
template <int NW>
__device__ int greedy_matches(const uint8_t *inner, int S, const uint32_t *outer_dw, int outer_ndw,
                              int T) {
  const int lane = (int)lane_id();
  const int h = (S + 1) >> 1;
  const int npos = S < 257 ? S : 257;
  uint32_t ch[NW];
  uint64_t fr[NW];
#pragma unroll
  for (int w = 0; w < NW; w++) {
    const int pos = lane + 64 * w;
    ch[w] = pos < npos ? (uint32_t)inner[pos] : 0x100u;
    const int lo = 64 * w;
    fr[w] = npos >= lo + 64 ? ~0ull : (npos > lo ? ((1ull << (npos - lo)) - 1ull) : 0ull);
  }
  int n_out = min(T, S + h);
  n_out = min(n_out, 257 + h);
  int m = 0;
  uint32_t od = 0;
  for (int i = 0; i < n_out; i++) {
    if ((i & 255) == 0) {
      const int dw = (i >> 2) + lane;
      od = dw < outer_ndw ? outer_dw[dw] : 0u;
    }
    const uint32_t c = ((uint32_t)__builtin_amdgcn_readlane((int)od, (i >> 2) & 63) >>
                        ((i & 3) << 3)) & 0xffu;
    const int lo = max(i - h, 0);
    int hi = min(i + h, S - 1);
    hi = min(hi, i < h ? 255 : 256);
    if (lo > hi) continue;
    bool found = false;
#pragma unroll
    for (int w = 0; w < NW; w++) {
      const int base = 64 * w;
      if (!found && hi >= base && lo <= base + 63) {
        const uint64_t eq = __ballot(ch[w] == c);
        const int a = max(lo - base, 0), b = min(hi - base, 63);
        const uint64_t win = (b == 63 ? ~0ull : ((2ull << b) - 1ull)) & ~((1ull << a) - 1ull);
        const uint64_t cand = eq & fr[w] & win;
        if (cand) {
          fr[w] ^= cand & (0ull - cand);
          m++;
          found = true;
        }
      }
    }
  }
  return m;
}

// Common prefix of a and b up to min(la, lb), one wave.
__device__ int common_prefix(const uint8_t *a, const uint8_t *b, int n) {
  const int lane = (int)lane_id();
  for (int base = 0; base < n; base += 64) {
    const int pos = base + lane;
    const bool diff = pos < n && a[pos] != b[pos];
    const uint64_t mask = __ballot(diff);
    if (mask) return base + __ffsll((unsigned long long)mask) - 1;
  }
  return n;
}

struct ExactArgs {
  const uint32_t *pairs;
  const unsigned long long *n_pairs;
  unsigned long long in_cap;
  DevColumn col_l, col_r;
  float threshold;
  int32_t slot;
  StageOut out;
  AuditOut audit;
  unsigned long long *jw_count;
};

__global__ __launch_bounds__(256) void k_exact(ExactArgs a) {
  const uint32_t lane = lane_id();
  const unsigned long long n_in = *a.n_pairs;
  const unsigned long long n_use = n_in < a.in_cap ? n_in : a.in_cap;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t n_waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const double thr = (double)a.threshold;
  unsigned long long evals = 0;
  for (uint64_t idx = wave; idx < n_use; idx += n_waves) {
    const uint32_t l = a.pairs[2 * idx], r = a.pairs[2 * idx + 1];
    const uint8_t *ca = a.col_l.bytes + (uint64_t)l * a.col_l.width;
    const uint8_t *cb = a.col_r.bytes + (uint64_t)r * a.col_r.width;
    const int la = a.col_l.len[l], lb = a.col_r.len[r];
    const int S = min(la, lb), T = max(la, lb);
    // device_func.cuh:71-79: the left cell stays inner unless strictly longer.
    const bool swap = la > lb;
    const uint8_t *inner = swap ? cb : ca;
    const uint8_t *outer = swap ? ca : cb;
    const uint32_t outer_width = swap ? a.col_l.width : a.col_r.width;
    const int p = common_prefix(ca, cb, S);
    int m = 0;
    if (S > 0) {
      const uint32_t *odw = reinterpret_cast<const uint32_t *>(outer);
      const int ndw = (int)(outer_width >> 2);
      if (S <= 64)
        m = greedy_matches<1>(inner, S, odw, ndw, T);
      else if (S <= 128)
        m = greedy_matches<2>(inner, S, odw, ndw, T);
      else
        m = greedy_matches<5>(inner, S, odw, ndw, T);
    }
    const float jaro = jaro_value(m, S, T);
    const double jw = winkler(jaro, p);
    evals++;
    if (lane == 0) {
      const bool pass = !(jw < thr);
      if (pass) {
        const unsigned long long pos = atomicAdd(a.out.count, 1ull);
        if (pos < a.out.cap) {
          a.out.pairs[2 * pos] = l;
          a.out.pairs[2 * pos + 1] = r;
        } else {
          *a.out.overflow = 1ull;
        }
      }
      if (fabs(jw - thr) < 1e-6) {
        const double jf = winkler_fused(jaro, p);
        const unsigned long long pos = atomicAdd(a.audit.count, 1ull);
        if (pos < a.audit.cap) {
          hb_audit_entry *e = reinterpret_cast<hb_audit_entry *>(a.audit.entries) + pos;
          e->l = l;
          e->r = r;
          e->slot = a.slot;
          e->flags = (((jw < thr) != (jf < thr)) ? 1 : 0) | (S >= 257 ? 2 : 0);
          e->jw = jw;
        } else {
          *a.out.overflow = 1ull;
        }
      }
    }
  }
  if (lane == 0 && evals) atomicAdd(a.jw_count, evals);
}

// ---- output merge -------------------------------------------------------------

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__global__ void k_merge_unique(const uint32_t *__restrict__ pairs,
                               const unsigned long long *__restrict__ n_pairs,
                               unsigned long long in_cap, unsigned long long *set,
                               unsigned long long set_cap, StageOut out) {
  const unsigned long long n = *n_pairs < in_cap ? *n_pairs : in_cap;
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x) {
    const uint32_t l = pairs[2 * i], r = pairs[2 * i + 1];
    const unsigned long long key = ((unsigned long long)l << 32) | r;
    unsigned long long slot = mix64(key) & (set_cap - 1);
    bool placed = false;
    for (unsigned long long probe = 0; probe < set_cap; probe++) {
      const unsigned long long prev = atomicCAS(set + slot, kEmpty, key);
      if (prev == kEmpty) {
        const unsigned long long pos = atomicAdd(out.count, 1ull);
        if (pos < out.cap) {
          out.pairs[2 * pos] = l;
          out.pairs[2 * pos + 1] = r;
        } else {
          *out.overflow = 1ull;
        }
        placed = true;
        break;
      }
      if (prev == key) {
        placed = true;
        break;
      }
      slot = (slot + 1) & (set_cap - 1);
    }
    if (!placed) *out.overflow = 1ull;
  }
}

__global__ void k_jw_from_counts(const int32_t *mstp, int64_t n, double *out, double *out_fused) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = mstp[4 * i], S = mstp[4 * i + 1], T = mstp[4 * i + 2], p = mstp[4 * i + 3];
  const float jaro = jaro_value(m, S, T);
  out[i] = winkler(jaro, p);
  out_fused[i] = winkler_fused(jaro, p);
}

inline unsigned grid_for(uint64_t n, unsigned block) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 0x7fffffffull) g = 0x7fffffffull;
  return (unsigned)g;
}

}  // namespace

void launch_build_sig(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint16_t *len,
                      SigRow *sig, hipStream_t stream) {
  if (n_rows <= 0) return;
  hipLaunchKernelGGL(k_build_sig, dim3(grid_for((uint64_t)n_rows, 256)), dim3(256), 0, stream,
                     bytes, width, n_rows, len, sig);
}

void launch_iota(uint32_t *out, uint32_t n, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256)), dim3(256), 0, stream, out, n);
}

void launch_fill_u64(unsigned long long *p, unsigned long long n, unsigned long long v,
                     hipStream_t stream) {
  if (!n) return;
  unsigned g = grid_for(n, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fill_u64, dim3(g), dim3(256), 0, stream, p, n, v);
}

void launch_run_flags(const uint32_t *sorted_keys, uint32_t n, uint32_t *flags,
                      hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(k_run_flags, dim3(grid_for(n, 256)), dim3(256), 0, stream, sorted_keys, n,
                     flags);
}

void launch_run_info(const uint32_t *sorted_l, const uint32_t *flags, const uint32_t *run_id,
                     uint32_t n_l, const uint32_t *sorted_r, uint32_t n_r, GroupArrays g,
                     hipStream_t stream) {
  hipLaunchKernelGGL(k_run_info, dim3(grid_for((uint64_t)n_l + 1, 256)), dim3(256), 0, stream,
                     sorted_l, flags, run_id, n_l, sorted_r, n_r, g);
}

void launch_single_group(uint32_t n_l, uint32_t n_r, GroupArrays g, hipStream_t stream) {
  hipLaunchKernelGGL(k_single_group, dim3(1), dim3(64), 0, stream, n_l, n_r, g);
}

void launch_filter_groups(GroupArrays g, const uint32_t *perm_l, const uint32_t *perm_r,
                          const SigRow *sig_l, const SigRow *sig_r, EqCheck eq, int has_sim,
                          const float coef[4], StageOut out, int n_cu, hipStream_t stream) {
  FilterArgs a;
  a.g = g;
  a.perm_l = perm_l;
  a.perm_r = perm_r;
  a.sig_l = sig_l;
  a.sig_r = sig_r;
  a.eq = eq;
  a.has_sim = has_sim;
  a.c0 = coef[0];
  a.c1 = coef[1];
  a.c2 = coef[2];
  a.c3 = coef[3];
  a.out = out;
  const unsigned grid = (unsigned)(n_cu > 0 ? n_cu : 256) * 4;
  hipLaunchKernelGGL(k_filter_tiles, dim3(grid), dim3(256), 0, stream, a);
  hipLaunchKernelGGL(k_filter_small, dim3(grid), dim3(256), 0, stream, a);
}

void launch_exact(const uint32_t *pairs, const unsigned long long *n_pairs,
                  unsigned long long in_cap, DevColumn col_l, DevColumn col_r, float threshold,
                  int32_t slot, StageOut out, AuditOut audit, unsigned long long *jw_count,
                  int n_cu, hipStream_t stream) {
  ExactArgs a;
  a.pairs = pairs;
  a.n_pairs = n_pairs;
  a.in_cap = in_cap;
  a.col_l = col_l;
  a.col_r = col_r;
  a.threshold = threshold;
  a.slot = slot;
  a.out = out;
  a.audit = audit;
  a.jw_count = jw_count;
  const unsigned grid = (unsigned)(n_cu > 0 ? n_cu : 256) * 8;
  hipLaunchKernelGGL(k_exact, dim3(grid), dim3(256), 0, stream, a);
}

void launch_merge_unique(const uint32_t *pairs, const unsigned long long *n_pairs,
                         unsigned long long in_cap, unsigned long long *set,
                         unsigned long long set_cap, StageOut out, int n_cu, hipStream_t stream) {
  const unsigned grid = (unsigned)(n_cu > 0 ? n_cu : 256) * 4;
  hipLaunchKernelGGL(k_merge_unique, dim3(grid), dim3(256), 0, stream, pairs, n_pairs, in_cap,
                     set, set_cap, out);
}

void launch_jw_from_counts(const int32_t *mstp, int64_t n, double *out, double *out_fused,
                           hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_jw_from_counts, dim3(grid_for((uint64_t)n, 256)), dim3(256), 0, stream,
                     mstp, n, out, out_fused);
}

}  // namespace hb
