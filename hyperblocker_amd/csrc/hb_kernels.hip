// gfx950 kernels of the rule-based blocking path.
//
// Reference semantics: core/gpu/global_func.cuh:104-185 (blocking_kernel),
// core/gpu/device_func.cuh:39-52 (equalities_kernel), :54-132
// (lev_jaro_ratio), :210-231 (lev_jaro_winkler_ratio).
//
// A batch of jobs runs as one sequence of launches over all of its job-rules
// (JRs, one rule of one job):
//   1. join (k_batch_join): every left run of equal first-equality code in a
//      JR's cached layout meets the right run of the same code (binary
//      search); a rule without equalities is one group covering L x R.
//   2. k_filter_tiles: pairs of every group are checked for
//      the remaining equalities (integer code compares) and an upper bound of
//      the first similarity predicate (byte-class histograms, v_sad_u8);
//      survivors are compacted with wave ballots into a shared pair list
//      tagged with their JR.  Large groups run as tiles (512 rows in lanes x
//      256 rows staged in LDS and read as broadcasts), small groups a wave
//      each, in the launch's last work-groups.
//   3. k_exact_two: each surviving pair (one per lane) gets the exact greedy
//      Jaro match count from bit-planes of its shorter string, then the
//      reference's fp32/fp64 tail, for two consecutive similarity
//      predicates per pass (in separate waves, decided in order).  Shorter
//      strings above 256 bytes go to k_exact (one wave per pair).  Later
//      predicates run as further passes on the shrinking list; pairs
//      passing a JR's last predicate enter the final list once per job
//      (hash set for jobs with several rules).
//   4. k_scatter: the final list is copied into per-job ranges.
//
// Floating point: contraction is off for this file, fp32 '/' is correctly
// rounded (hipcc default), so every operation of the Jaro/Winkler tail
// rounds once, in the reference's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "../../include/hyperblocker_amd.h"
#include "hb_internal.h"
#define HB_ALLOW_FN __device__ __forceinline__
#include "hb_allow.h"

#pragma clang fp contract(off)

namespace hb {
namespace {

constexpr unsigned long long kEmpty = ~0ull;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Values the whole wave holds (loaded from uniform addresses, which the
// compiler cannot prove uniform), moved to scalar registers.  The builtin
// returns a signed int: each half is taken back to uint32_t before the
// halves are joined (an int low half would sign-extend into the high one).
__device__ __forceinline__ uint32_t wave_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ uint64_t wave_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ float wave_f32(float v) {
  return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}
template <class T>
__device__ __forceinline__ T *wave_ptr(T *p) {
  return (T *)wave_u64((uint64_t)p);
}

__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int lane) {
  uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}

// Appends the entry of every lane with ok set to a pair list; one atomic
// per wave.  A full list raises the batch's overflow flag (the batch is then
// run again with larger lists).
__device__ __forceinline__ void list_append(bool ok, const PairEntry &e, const PairList &out,
                                            unsigned long long *overflow) {
  const uint64_t mask = __ballot(ok);
  if (mask == 0) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(out.count, (unsigned long long)__popcll(mask));
  base = readlane64(base, leader);
  if (ok) {
    const unsigned long long pos = base + rank_in(mask);
    if (pos < out.cap)
      out.e[pos] = e;
    else
      *overflow = 1ull;
  }
}

// Adds one for every lane with ok set to counter k of the lane's job: one
// atomic per distinct job in the wave (usually one).
__device__ __forceinline__ void job_count(const BatchDev &b, int k, bool ok, uint32_t job) {
  uint64_t mask = __ballot(ok);
  while (mask) {
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t lj = (uint32_t)__builtin_amdgcn_readlane((int)job, leader);
    const uint64_t same = __ballot(ok && job == lj);
    if ((int)lane_id() == leader)
      atomicAdd(b.job_cnt + (size_t)job_row(k, blockIdx.x & (kStatStripes - 1)) * b.cnt_stride + lj,
                (unsigned long long)__popcll(same));
    mask &= ~same;
  }
}

// Adds v of every lane with ok set to counter k of the lane's job: one
// atomic per distinct job in the wave.  Every lane of the wave must call it.
__device__ __forceinline__ void job_add(const BatchDev &b, int k, bool ok, uint32_t job,
                                        unsigned long long v) {
  uint64_t mask = __ballot(ok);
  while (mask) {
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t lj = (uint32_t)__builtin_amdgcn_readlane((int)job, leader);
    const bool same = ok && job == lj;
    unsigned long long sum = same ? v : 0ull;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d);
    if ((int)lane_id() == leader)
      atomicAdd(b.job_cnt + (size_t)job_row(k, blockIdx.x & (kStatStripes - 1)) * b.cnt_stride + lj,
                sum);
    mask &= ~__ballot(same);
  }
}

// A slot of the lane's job among fill[job]: consecutive per job, one atomic
// per distinct job in the wave.
__device__ __forceinline__ unsigned long long job_slot(bool ok, uint32_t job,
                                                       unsigned long long *fill) {
  unsigned long long mine = 0;
  uint64_t mask = __ballot(ok);
  while (mask) {
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t lj = (uint32_t)__builtin_amdgcn_readlane((int)job, leader);
    const uint64_t same = __ballot(ok && job == lj);
    unsigned long long base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(fill + lj, (unsigned long long)__popcll(same));
    base = readlane64(base, leader);
    if (ok && job == lj) mine = base + rank_in(same);
    mask &= ~same;
  }
  return mine;
}

// Survivor buffer of one wave in LDS, for kernels whose waves run
// independent loops (no work-group barrier): entries accumulate under a
// wave-uniform count held in a register, and the wave moves them to the
// global list with one atomic when the buffer would overflow, and at the end.
// LDS accesses of one wave complete in order; the wave barriers keep the
// compiler from moving them across the hand-over between lanes.
constexpr uint32_t kWaveOut = 128;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void wave_flush(PairEntry *buf, uint32_t &n, const PairList &out,
                                           unsigned long long *overflow) {
  if (n == 0) return;
  wave_sync();
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(out.count, (unsigned long long)n);
  base = readlane64(base, 0);
  for (uint32_t i = lane_id(); i < n; i += 64) {
    const unsigned long long pos = base + i;
    if (pos < out.cap)
      out.e[pos] = buf[i];
    else
      *overflow = 1ull;
  }
  wave_sync();
  n = 0;
}

__device__ __forceinline__ void wave_append(bool ok, const PairEntry &e, PairEntry *buf,
                                            uint32_t &n, const PairList &out,
                                            unsigned long long *overflow) {
  const uint64_t mask = __ballot(ok);
  if (mask == 0) return;
  const uint32_t c = (uint32_t)__popcll(mask);
  if (n + c > kWaveOut) wave_flush(buf, n, out, overflow);
  if (ok) buf[n + rank_in(mask)] = e;
  n += c;
}

// Survivor buffer of a filter work item: 4-byte entries a | b << 10, with a
// the A row's index in the item's A tile (< 1024) and b the B row's index
// from the start of the item's B segment (< 65536).  The group flushes it
// at the end of every work item (and at a chunk boundary when half full),
// translating entries to table rows.  Entries past the buffer go to the
// global list directly, one atomic per wave.
constexpr uint32_t kTileOut = 2048;
static_assert(256 * 4 <= kTileOut / 2, "one filter step must fit half the tile buffer");
struct TileOut {
  uint32_t e[kTileOut];
  uint32_t n;
  uint32_t dense;  // the last flush held many entries: flush between segments of B rows too
  unsigned long long base;
};

// What a TileOut entry needs to become a pair list entry.
struct TileCtx {
  const uint32_t *perm_a;  // join-order permutation of the A side, from the group's first row
  const uint32_t *perm_b;  // same for the B side, from the item's first B row
  uint32_t a_base;         // A tile's first index in the group
  uint32_t jr;
  bool a_is_l;
};

__device__ __forceinline__ PairEntry tile_entry(const TileCtx &c, uint32_t a, uint32_t b) {
  const uint32_t ar = c.perm_a[c.a_base + a], br = c.perm_b[b];
  return PairEntry{c.a_is_l ? ar : br, c.a_is_l ? br : ar, c.jr, 0u};
}

__device__ __forceinline__ void tile_append(bool ok, uint32_t a, uint32_t b, const TileCtx &c,
                                            TileOut &to, const PairList &out,
                                            unsigned long long *overflow) {
  const uint64_t mask = __ballot(ok);
  if (mask == 0) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(&to.n, (uint32_t)__popcll(mask));
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
  const uint32_t pos = base + rank_in(mask);
  if (ok && pos < kTileOut) to.e[pos] = a | (b << 10);
  const bool spill = ok && pos >= kTileOut;
  if (__builtin_amdgcn_ballot_w64(spill))
    list_append(spill, spill ? tile_entry(c, a, b) : PairEntry{}, out, overflow);
}

// Called by every thread of the group.  Leaves to.n == 0.
__device__ __forceinline__ void tile_flush(TileOut &to, const TileCtx &c, const PairList &out,
                                           unsigned long long *overflow) {
  __syncthreads();
  const uint32_t n = min(to.n, kTileOut);
  if (threadIdx.x == 0) {
    to.dense = to.n >= kTileOut / 4 ? 1u : 0u;
    to.base = n ? atomicAdd(out.count, (unsigned long long)n) : 0ull;
  }
  __syncthreads();
  const unsigned long long base = to.base;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned long long pos = base + i;
    const uint32_t v = to.e[i];
    if (pos < out.cap)
      out.e[pos] = tile_entry(c, v & 1023u, v >> 10);
    else
      *overflow = 1ull;
  }
  __syncthreads();
  if (threadIdx.x == 0) to.n = 0;
  __syncthreads();
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
  s.rlen = (L >= 1u && L <= 255u) ? 1.0f / (float)L : INFINITY;
  s.pad = 0;
  sig_out[r] = s;
}

// The counter class of a non-zero byte (kHistWords layout): byte & 127,
// with classes 0-31 counted in class + 32.  Its low 5 bits are the byte's,
// so the signatures' 32 classes are sums of these.
__device__ __forceinline__ uint32_t byte_class(uint32_t c) {
  const uint32_t k = c & 127u;
  return k < 32u ? k + 32u : k;
}

// Two words of u8 class counters (classes 8k .. 8k + 3 and 8k + 4 .. 8k + 7)
// as one word of nibble counters saturated at 15 (the kHistWords layout).
__device__ __forceinline__ uint32_t sat_nibbles(uint32_t lo4, uint32_t hi4) {
  auto sat = [](uint32_t x) {
    const uint32_t big = (((x >> 4) & 0x0f0f0f0fu) + 0x0f0f0f0fu) & 0x10101010u;  // byte >= 16
    return (x | ((big >> 4) * 0x0fu)) & 0x0f0f0f0fu;
  };
  return sat(lo4) | (sat(hi4) << 4);
}

// 128-class counters of each cell (layout: kHistWords).  Each thread
// counts one row in its own column of an LDS array, so no synchronisation
// is needed.
__global__ __launch_bounds__(256) void k_build_hist(const uint8_t *__restrict__ bytes,
                                                   uint32_t width, int64_t n,
                                                   uint32_t *__restrict__ hist) {
  __shared__ uint32_t s_h[kClassWords][256];
  const uint32_t t = threadIdx.x;
  const int64_t r = (int64_t)blockIdx.x * 256 + t;
#pragma unroll
  for (int w = 0; w < kClassWords; w++) s_h[w][t] = 0;
  if (r >= n) return;
  const uint4 *cell = reinterpret_cast<const uint4 *>(bytes + r * (int64_t)width);
  const uint32_t n16 = width >> 4;
  bool end = false;
  uint32_t L = 0, pre = 0;
  for (uint32_t k = 0; k < n16 && !end; k++) {
    const uint4 v = cell[k];
    if (k == 0) pre = v.x;
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const uint32_t c = (w4[q] >> (8 * b)) & 0xffu;
        end |= (c == 0);
        L += !end;
        if (!end) s_h[byte_class(c) >> 2][t] += 1u << ((c & 3u) * 8u);
      }
    }
  }
  uint4 *dst = reinterpret_cast<uint4 *>(hist + r * kHistWords);
  dst[0] = make_uint4(pre, L | (L > 255u ? kLongFlag : 0u), 0u, 0u);
#pragma unroll
  for (int w = 1; w < kHistWords / 4; w++)
    dst[w] = make_uint4(sat_nibbles(s_h[8 * w][t], s_h[8 * w + 1][t]),
                        sat_nibbles(s_h[8 * w + 2][t], s_h[8 * w + 3][t]),
                        sat_nibbles(s_h[8 * w + 4][t], s_h[8 * w + 5][t]),
                        sat_nibbles(s_h[8 * w + 6][t], s_h[8 * w + 7][t]));
}

__global__ void k_build_planes(const uint8_t *__restrict__ bytes, uint32_t width, int64_t n,
                               uint32_t nwc, uint32_t *__restrict__ planes) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint8_t *cell = bytes + r * (int64_t)width;
  uint32_t *dst = planes + r * (int64_t)nwc * 8;
  for (uint32_t w = 0; w < nwc; w++) {
    uint32_t P[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = 0; b < 32; b++) {
      const uint32_t pos = 32 * w + b;
      const uint32_t c = pos < width ? cell[pos] : 0u;
#pragma unroll
      for (int k = 0; k < 8; k++) P[k] |= ((c >> k) & 1u) << b;
    }
    uint4 *d4 = reinterpret_cast<uint4 *>(dst + 8 * w);
    d4[0] = make_uint4(P[0], P[1], P[2], P[3]);
    d4[1] = make_uint4(P[4], P[5], P[6], P[7]);
  }
}

// 8 x 8 bit transpose (Hacker's Delight, transpose8): bit j of byte i moves
// to bit i of byte j.  For 8 consecutive cell bytes, byte k of the result
// holds bit k of each of them: 8 bits of bit-plane k.
__device__ __forceinline__ uint64_t transpose8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

// Number of non-zero bytes of a dword.
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t x) {
  const uint32_t y = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;  // high bit set where the low 7 bits are not 0
  return (uint32_t)__popc((y | x) & 0x80808080u);
}

// Adds the 32-class counts of one 32-byte block to the signature counters
// h (counter b in byte b & 3 of word b >> 2), from the block's bit-planes:
// the bytes of class k (byte & 31 == k) are the AND over bits b < 5 of
// (P[b] if bit b of k is set, else ~P[b]), built as a tree over bits 4..0
// (62 two-input operations), counted with one popcount each.  Class 0 also
// needs a set bit among bits 5-7, so zero bytes (the NUL padding) count
// nowhere.
__device__ __forceinline__ void count_classes(const uint32_t (&P)[8], uint32_t (&h)[8]) {
  uint32_t t1[2] = {~P[4], P[4]};
  uint32_t t2[4], t3[8], t4[16];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    t2[2 * i] = t1[i] & ~P[3];
    t2[2 * i + 1] = t1[i] & P[3];
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    t3[2 * i] = t2[i] & ~P[2];
    t3[2 * i + 1] = t2[i] & P[2];
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    t4[2 * i] = t3[i] & ~P[1];
    t4[2 * i + 1] = t3[i] & P[1];
  }
  // t4[i]: bits 4..1 of the byte equal i's bits 3..0 (i = class >> 1)
#pragma unroll
  for (int i = 0; i < 16; i++) {
    uint32_t m0 = t4[i] & ~P[0];
    if (i == 0) m0 &= P[5] | P[6] | P[7];  // class 0 of a non-zero byte
    const uint32_t m1 = t4[i] & P[0];
    const int k0 = 2 * i, k1 = 2 * i + 1;  // classes; counter k in byte k & 3 of word k >> 2
    h[k0 >> 2] += (uint32_t)__popc(m0) << (8 * (k0 & 3));
    h[k1 >> 2] += (uint32_t)__popc(m1) << (8 * (k1 & 3));
  }
}

// Every per-row structure of a similarity column from one read of its
// cells, four lanes per row: lane q takes the row's 32-byte blocks q, q + 4,
// ...  Bit-planes of a block: four 8 x 8 bit transposes.  Without refine
// rows (HIST false) the signature's 32 counters are counted from the
// block's bit-planes in registers and summed over the row's four lanes;
// with them, one LDS atomic add per byte goes into the row's 128 counters
// (byte_class, u8 each; zero bytes, which follow the cell's NUL, add
// nothing), from which the signature's 32 counters (byte & 31) are sums of
// four.  The length counts the non-zero bytes (every byte after the NUL is
// zero).  Results equal those of k_build_sig, k_build_hist and
// k_build_planes.
template <bool HIST>
__global__ __launch_bounds__(256) void k_build_col(const uint8_t *__restrict__ bytes,
                                                  uint32_t width, int64_t n, uint32_t nwc,
                                                  uint16_t *__restrict__ len_out,
                                                  SigRow *__restrict__ sig_out,
                                                  uint32_t *__restrict__ hist,
                                                  uint32_t *__restrict__ planes) {
  __shared__ uint32_t s_h[HIST ? kClassWords : 1][64];  // [class word][row of the work-group]
  const uint32_t t = threadIdx.x, q = t & 3u, rl = t >> 2;
  const int64_t r = (int64_t)blockIdx.x * 64 + rl;
  const bool valid = r < n;
  if (HIST) {
    for (uint32_t i = t; i < kClassWords * 64; i += 256) (&s_h[0][0])[i] = 0;
    __syncthreads();
  }
  uint32_t L = 0, pre = 0;
  uint32_t h[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // HIST false: this lane's blocks' counters
  if (valid) {
    const uint8_t *cell = bytes + r * (int64_t)width;
    uint32_t *pl = planes + r * (int64_t)nwc * 8;
    const uint32_t n32 = (width + 31u) >> 5;
    for (uint32_t blk = q; blk < n32; blk += 4) {
      const uint4 *src = reinterpret_cast<const uint4 *>(cell + 32 * blk);
      const uint4 a = src[0];
      const uint4 b = 32 * blk + 16 < width ? src[1] : make_uint4(0u, 0u, 0u, 0u);
      if (blk == 0) pre = a.x;
      const uint32_t w8[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int d = 0; d < 8; d++) {
        L += nonzero_bytes(w8[d]);
        if (HIST) {
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const uint32_t c = (w8[d] >> (8 * e)) & 0xffu;
            const uint32_t inc = c ? 1u << ((c & 3u) * 8u) : 0u;
            atomicAdd(&s_h[byte_class(c) >> 2][rl], inc);
          }
        }
      }
      if (blk < nwc || !HIST) {
        uint64_t X[4];
#pragma unroll
        for (int g = 0; g < 4; g++) X[g] = transpose8(((uint64_t)w8[2 * g + 1] << 32) | w8[2 * g]);
        uint32_t P[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const uint32_t sh = 8 * k;
          P[k] = (uint32_t)((X[0] >> sh) & 0xffu) | ((uint32_t)((X[1] >> sh) & 0xffu) << 8) |
                 ((uint32_t)((X[2] >> sh) & 0xffu) << 16) | ((uint32_t)((X[3] >> sh) & 0xffu) << 24);
        }
        if (!HIST) count_classes(P, h);
        if (blk < nwc) {
          uint4 *d4 = reinterpret_cast<uint4 *>(pl + 8 * blk);
          d4[0] = make_uint4(P[0], P[1], P[2], P[3]);
          d4[1] = make_uint4(P[4], P[5], P[6], P[7]);
        }
      }
    }
  }
  // sums over the row's four lanes (adjacent in the wave)
#pragma unroll
  for (int m = 1; m <= 2; m <<= 1) {
    L += (uint32_t)__shfl_xor((int)L, m);
    pre |= (uint32_t)__shfl_xor((int)pre, m);  // only lane 0 of the row holds it
    if (!HIST) {
#pragma unroll
      for (int j = 0; j < 8; j++) h[j] += (uint32_t)__shfl_xor((int)h[j], m);
    }
  }
  if (HIST) __syncthreads();
  if (!valid) return;
  if (HIST) {  // lane q writes nibble words 4q .. 4q + 3 of the refine row (classes
               // 32q .. 32q + 31; lane 0's are never counted, so it writes the prefix and length)
    uint32_t o[8];
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = s_h[8 * q + j][rl];
    uint4 *dst = reinterpret_cast<uint4 *>(hist + r * kHistWords + 4 * q);
    dst[0] = q == 0 ? make_uint4(pre, L | (L > 255u ? kLongFlag : 0u), 0u, 0u)
                    : make_uint4(sat_nibbles(o[0], o[1]), sat_nibbles(o[2], o[3]),
                                 sat_nibbles(o[4], o[5]), sat_nibbles(o[6], o[7]));
  }
  if (q == 0) {
    len_out[r] = (uint16_t)min(L, 65535u);
    SigRow sg;
#pragma unroll
    for (int j = 0; j < 8; j++)  // classes 4j .. 4j+3 of 32 = four of the 128 each
      sg.h[j] = HIST ? s_h[j][rl] + s_h[j + 8][rl] + s_h[j + 16][rl] + s_h[j + 24][rl] : h[j];
    sg.pre = pre;
    sg.len = L | (L > 255u ? kLongFlag : 0u);
    sg.rlen = (L >= 1u && L <= 255u) ? 1.0f / (float)L : INFINITY;
    sg.pad = 0;
    sig_out[r] = sg;
  }
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

// ---- device dictionary ---------------------------------------------------

// 64-bit hash of a cell's 16-byte chunks (NUL padded, so equal strings give
// equal chunks): multiply-xor per 8-byte word, then MurmurHash3's 64-bit
// finalizer over the hash and the length.
__device__ __forceinline__ uint64_t dict_mix(uint64_t h, uint64_t w) {
  h = (h ^ w) * 0xff51afd7ed558ccdull;
  return h ^ (h >> 32);
}
__device__ __forceinline__ uint64_t dict_final(uint64_t h, uint32_t len) {
  h ^= (uint64_t)len * 0x9E3779B97F4A7C15ull;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

// Position of the first zero byte of a dword (4 if none).
__device__ __forceinline__ uint32_t first_zero_byte(uint32_t x) {
  const uint32_t z = (x - 0x01010101u) & ~x & 0x80808080u;
  return z ? (uint32_t)(__builtin_ctz(z) >> 3) : 4u;
}

// Length and number of 16-byte chunks of a cell (the chunks hold the cell's
// bytes and, in NUL mode, its terminating NUL), and its hash.
__device__ __forceinline__ uint64_t dict_hash_cell(const uint4 *cell, uint32_t width,
                                                   uint32_t fixed_len, uint32_t hash_bits,
                                                   uint32_t &len, uint32_t &chunks) {
  uint64_t h = 0x9E3779B97F4A7C15ull;
  const uint32_t n16 = width >> 4;
  len = fixed_len;
  chunks = fixed_len ? (fixed_len + 15) >> 4 : n16;
  for (uint32_t k = 0; k < chunks; k++) {
    const uint4 v = cell[k];
    h = dict_mix(h, ((uint64_t)v.y << 32) | v.x);
    h = dict_mix(h, ((uint64_t)v.w << 32) | v.z);
    if (!fixed_len) {
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      uint32_t pos = 16;
#pragma unroll
      for (int q = 3; q >= 0; q--) {
        const uint32_t z = first_zero_byte(w[q]);
        if (z < 4) pos = 4 * q + z;
      }
      if (pos < 16) {
        len = 16 * k + pos;
        chunks = k + 1;
        break;
      }
    }
  }
  h = dict_final(h, len);
  if (hash_bits < 64) h &= (1ull << hash_bits) - 1ull;
  return h == kDictEmpty ? kDictEmpty - 1ull : h;
}

// One thread per row: find or claim the slot of the cell's hash; `code`
// holds the slot until k_dict_verify replaces it by the value's code.
__global__ void k_dict_insert(const uint8_t *__restrict__ cells, uint32_t width,
                              uint32_t fixed_len, uint32_t n, DictDev d,
                              uint32_t *__restrict__ code, unsigned long long *fb_count) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) *fb_count = 0ull;  // k_dict_verify counts after this kernel
  if (r >= n) return;
  const uint4 *cell = reinterpret_cast<const uint4 *>(cells + (size_t)r * width);
  uint32_t len, chunks;
  const uint64_t h = dict_hash_cell(cell, width, fixed_len, d.hash_bits, len, chunks);
  uint32_t slot = (uint32_t)(h ^ (h >> 31)) & d.mask;
  uint32_t c = kNoCode;
  for (uint32_t p = 0; p < d.max_probe; p++) {
    unsigned long long k = __hip_atomic_load(d.keys + slot, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    if (k == kDictEmpty) {
      k = atomicCAS(d.keys + slot, kDictEmpty, (unsigned long long)h);
      if (k == kDictEmpty) {  // this row's cell becomes the slot's representative
        const unsigned long long bytes = (unsigned long long)chunks * 16;
        const unsigned long long off = atomicAdd(d.arena_used, bytes);
        unsigned long long rep = kDictNoRep;
        if (off + bytes <= d.arena_cap && chunks < 65536) {
          uint4 *dst = reinterpret_cast<uint4 *>(d.arena + off);
          for (uint32_t q = 0; q < chunks; q++) dst[q] = cell[q];
          rep = (off << 16) | chunks;
        }
        d.reps[slot] = rep;
        d.slot_code[slot] = atomicAdd(d.n_codes, 1u);
        c = slot;
        break;
      }
    }
    if (k == h) {
      c = slot;
      break;
    }
    slot = (slot + 1) & d.mask;
  }
  code[r] = c;
}

// One thread per row: the cell must equal its slot's representative; rows
// that do not (or have no slot) are listed for the host.
__device__ __forceinline__ void dict_verify_row(const uint8_t *__restrict__ cells, uint32_t width,
                                                uint32_t fixed_len, uint32_t r, const DictDev &d,
                                                uint32_t *__restrict__ code,
                                                uint32_t *__restrict__ fb_rows,
                                                unsigned long long *fb_count) {
  const uint32_t c = code[r];
  bool ok = c != kNoCode;
  if (ok) {
    const unsigned long long rep = d.reps[c];
    const uint4 *cell = reinterpret_cast<const uint4 *>(cells + (size_t)r * width);
    uint32_t chunks = fixed_len ? (fixed_len + 15) >> 4 : 0;
    if (!fixed_len) {  // the cell's chunks up to the one holding its NUL
      for (uint32_t k = 0; k < (width >> 4); k++) {
        const uint4 v = cell[k];
        chunks = k + 1;
        if (first_zero_byte(v.x) < 4 || first_zero_byte(v.y) < 4 || first_zero_byte(v.z) < 4 ||
            first_zero_byte(v.w) < 4)
          break;
      }
    }
    ok = rep != kDictNoRep && (uint32_t)(rep & 0xffffu) == chunks;
    if (ok) {
      const uint4 *src = reinterpret_cast<const uint4 *>(d.arena + (rep >> 16));
      for (uint32_t q = 0; q < chunks && ok; q++) {
        const uint4 a = cell[q], b = src[q];
        ok = a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
      }
    }
  }
  if (ok) {
    code[r] = d.slot_code[c];
  } else {
    code[r] = kNoCode;
    fb_rows[atomicAdd(fb_count, 1ull)] = r;
  }
}

__global__ __launch_bounds__(1024) void k_dict_verify(const uint8_t *__restrict__ cells,
                                                     uint32_t width, uint32_t fixed_len,
                                                     uint32_t n, DictDev d,
                                                     uint32_t *__restrict__ code,
                                                     uint32_t *__restrict__ fb_rows,
                                                     unsigned long long *fb_count,
                                                     unsigned long long *status) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dict_verify_row(cells, width, fixed_len, r, d, code, fb_rows, fb_count);
  // The last work-group to finish reports (rows left to the host, codes
  // handed out): every group's counts are in once it has seen them all done.
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(d.done, 1u) == gridDim.x - 1) {
      __threadfence();
      status[0] = __hip_atomic_load(fb_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      status[1] = __hip_atomic_load(d.n_codes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *d.done = 0u;  // for the next coding (codings run in order on one stream)
    }
  }
}

// Key cells of a composite key: 0xC0 + k, then the k codes.
__global__ void k_compose_keys(KeyColumns cols, uint32_t n, uint8_t *__restrict__ cells,
                               uint32_t width) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t *dst = reinterpret_cast<uint32_t *>(cells + (size_t)r * width);
  // byte 0 = marker, bytes 1 + 4q .. 4 + 4q = code q: assembled as dwords
  uint32_t carry = 0xC0u + cols.k;
  uint32_t w = 0;
  for (uint32_t q = 0; q < cols.k; q++) {
    const uint32_t v = cols.code[q][r];
    dst[w++] = carry | (v << 8);
    carry = v >> 24;
  }
  dst[w++] = carry;
  for (; w < (width >> 2); w++) dst[w] = 0u;
}

// Composite codes of rows with a column code still left to the host
// (kNoCode) become kNoCode: the layouts built from them are flagged bad.
__global__ void k_mask_keys(KeyColumns cols, uint32_t n, uint32_t *__restrict__ code) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  bool pending = false;
  for (uint32_t q = 0; q < cols.k; q++) pending |= cols.code[q][r] == kNoCode;
  if (pending) code[r] = kNoCode;
}

// A work-group unpacks blocks of rows_per_group consecutive rows (grid
// stride).  A block is one contiguous span of the input: it is loaded into
// LDS (16-byte loads when the rows are 16-byte aligned), each cell's length
// (up to its first NUL, at most `size`) is found there, and each column's
// cells of the block, one contiguous span of that column, are written in
// 16-byte chunks with every byte from the cell's end on zero.
template <bool ALIGNED16>
__global__ __launch_bounds__(256) void k_unpack_rows(const uint8_t *__restrict__ rows,
                                                    uint64_t row_bytes, uint32_t n_rows,
                                                    const UnpackColumns inl,
                                                    const UnpackColumn *__restrict__ cols_dev,
                                                    uint32_t n_cols, uint32_t rows_per_group) {
  const UnpackColumn *__restrict__ cols = n_cols <= kUnpackArgCols ? inl.c : cols_dev;
  extern __shared__ uint4 s_mem4[];
  uint8_t *s_rows = reinterpret_cast<uint8_t *>(s_mem4);
  const uint32_t *s_rows32 = reinterpret_cast<const uint32_t *>(s_mem4);
  const uint32_t span_max = (uint32_t)(rows_per_group * row_bytes);
  // staged rows (plus 32 bytes read past the end); one bit per staged byte,
  // set where the byte is zero; the length of every cell of the block
  const uint32_t rows_room = ((span_max + 31u) & ~31u) + 32u;
  uint32_t *s_zero = reinterpret_cast<uint32_t *>(s_rows + rows_room);
  uint16_t *s_len = reinterpret_cast<uint16_t *>(s_zero + (span_max + 31u) / 32u + 1u);
  const uint32_t tid = threadIdx.x;
  // s_keep[k]: a mask of the first k bytes of a 16-byte chunk (k <= 16)
  __shared__ uint4 s_keep[17];
  if (tid < 17) {
    const int k = (int)tid;
    auto dw = [k](int d) -> uint32_t {
      const int nb = k - 4 * d;
      return nb >= 4 ? ~0u : (nb <= 0 ? 0u : (1u << (8 * nb)) - 1u);
    };
    s_keep[tid] = make_uint4(dw(0), dw(1), dw(2), dw(3));
  }
  const uint32_t n_blocks = (n_rows + rows_per_group - 1) / rows_per_group;
  for (uint32_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const uint32_t r0 = blk * rows_per_group;
    const uint32_t nr = min(rows_per_group, n_rows - r0);
    const uint32_t span = (uint32_t)(nr * row_bytes);
    const uint8_t *src = rows + (size_t)r0 * row_bytes;
    __syncthreads();  // the previous block's reads of LDS are done
    // Zero-byte bitmap, 16 bytes (one 16-bit piece) at a time.  The flags of
    // a dword's zero bytes sit at bits 7, 15, 23 and 31 of m; one multiply
    // gathers them into bits 21-24 (each lands on its own bit, so nothing
    // carries).  The pieces up to the bitmap word's end past `span` are
    // zero bytes.
    uint16_t *s_zero16 = reinterpret_cast<uint16_t *>(s_zero);
    const uint32_t n16_words = ((span + 31u) / 32u) * 2u;
    auto zero_bits = [](const uint4 x) -> uint32_t {
      const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
      uint32_t z = 0;
#pragma unroll
      for (int h = 0; h < 4; h++) {
        const uint32_t v = xs[h];
        const uint32_t m = ~(((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;  // zero bytes
        z |= (((m >> 7) * 0x00204081u) >> 21 & 0xfu) << (4 * h);
      }
      return z;
    };
    if (ALIGNED16) {  // each 16 bytes' piece of the bitmap as they are loaded
      const uint4 *src4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll 8
      for (uint32_t i = tid; i < span / 16; i += 256) {
        const uint4 x = src4[i];
        s_mem4[i] = x;
        s_zero16[i] = (uint16_t)zero_bits(x);
      }
      for (uint32_t i = span / 16 + tid; i < n16_words; i += 256) s_zero16[i] = 0xffffu;
    } else {
      for (uint32_t i = tid; i < span; i += 256) s_rows[i] = src[i];
    }
    // the bytes from span to the next 32-byte boundary read as zero
    for (uint32_t i = span + tid; i < ((span + 31u) & ~31u) + 32u; i += 256) s_rows[i] = 0;
    __syncthreads();
    if (!ALIGNED16) {
      for (uint32_t i = tid; i < n16_words; i += 256) s_zero16[i] = (uint16_t)zero_bits(s_mem4[i]);
      __syncthreads();
    }
    // each cell's length: its first zero byte, at most `size`
    for (uint32_t i = tid; i < nr * n_cols; i += 256) {
      const uint32_t j = i / n_cols, c = i - j * n_cols;
      const uint32_t start = j * (uint32_t)row_bytes + (uint32_t)cols[c].offset;
      const uint32_t size = cols[c].size;
      uint32_t len = size;
      for (uint32_t w = start >> 5; (w << 5) < start + size; w++) {
        uint32_t z = s_zero[w];
        if ((w << 5) < start) z &= ~0u << (start & 31);
        if (z) {
          len = min(size, (w << 5) + (uint32_t)__builtin_ctz(z) - start);
          break;
        }
      }
      s_len[i] = (uint16_t)len;
    }
    __syncthreads();
    for (uint32_t c = 0; c < n_cols; c++) {
      const uint32_t cpc = cols[c].width >> 4;  // 16-byte chunks per cell
      const uint32_t off = (uint32_t)cols[c].offset;
      uint4 *dst = reinterpret_cast<uint4 *>(cols[c].out + (size_t)r0 * cols[c].width);
      // thread t takes chunk q of cells j, j + per, ... (one division per column)
      const uint32_t per = cpc <= 256u ? 256u / cpc : 0u;
      const uint32_t tj = per ? tid / cpc : 0u, tq = per ? tid - tj * cpc : 0u;
      for (uint32_t it = 0;; it++) {
        uint32_t j, q;
        if (per) {
          if (tj >= per) break;
          j = tj + it * per;
          q = tq;
          if (j >= nr) break;
        } else {  // cells wider than 4 KiB
          const uint32_t i = tid + it * 256u;
          if (i >= nr * cpc) break;
          j = i / cpc;
          q = i - j * cpc;
        }
        const uint32_t len = s_len[j * n_cols + c];
        const uint32_t base = 16u * q;
        // 16 bytes from any byte position: five dwords and four byte shifts
        // (a chunk past the cell's end reads bytes of the block, all masked)
        const uint32_t pos = j * (uint32_t)row_bytes + off + base;
        const uint32_t w = pos >> 2, sh = pos & 3u;
        const uint32_t d0 = s_rows32[w], d1 = s_rows32[w + 1], d2 = s_rows32[w + 2],
                       d3 = s_rows32[w + 3], d4 = s_rows32[w + 4];
        // the bytes of this chunk inside the cell: min(16, len - base), or 0
        const uint4 m = s_keep[len > base ? min(len - base, 16u) : 0u];
        uint4 v;
        v.x = __builtin_amdgcn_alignbyte(d1, d0, sh) & m.x;
        v.y = __builtin_amdgcn_alignbyte(d2, d1, sh) & m.y;
        v.z = __builtin_amdgcn_alignbyte(d3, d2, sh) & m.z;
        v.w = __builtin_amdgcn_alignbyte(d4, d3, sh) & m.w;
        dst[j * cpc + q] = v;
      }
    }
  }
}

// Join-order keys: (dictionary code << 8) | min(length, 255) of the rule's
// first similarity column (0 without one), or the length alone when the rule
// has no equality.  Sorting them groups rows by code and orders each group by
// length.
__device__ __forceinline__ uint64_t checked_code(const uint32_t *code, uint32_t i,
                                                 uint64_t bound, uint32_t *bad) {
  const uint64_t c = code ? (uint64_t)code[i] : 0ull;
  if (c < bound) return c;
  *bad = 1u;
  return 0ull;
}

__global__ void k_layout_keys(const uint32_t *__restrict__ code, const uint16_t *__restrict__ len,
                              uint32_t n, uint64_t bound, uint32_t *bad,
                              uint64_t *__restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = checked_code(code, i, bound, bad);
  const uint32_t l = len ? min((uint32_t)len[i], 255u) : 0u;
  keys[i] = (c << 8) | l;
}

// ---- join layouts by counting sort (keys of few codes) ------------------
//
// Buckets (code, min(length, 255)) for the join order and (code, first
// byte) for the first-byte order, n_codes * 256 of each.  One counting
// launch takes each row's rank in its two buckets (the order within a
// bucket is arbitrary: its rows share the key the layout orders by); its
// last work-group to finish turns the counts into bucket offsets and the
// runs of equal code and leaves the counters zero again; k_lsmall_scatter
// places every row and its signature.

__device__ __forceinline__ uint32_t lsmall_code(const LayoutSmall &a, uint32_t i, uint32_t *bad) {
  const uint32_t c = a.code ? a.code[i] : 0u;
  if (c < a.n_codes) return c;
  if (bad) *bad = 1u;
  return 0u;
}

// Whether this work-group is the last of the launch to get here: each
// group's earlier global writes and atomics are visible to the last one
// (every group fences before counting itself done, the last fences after).
__device__ __forceinline__ bool last_group(uint32_t *done) {
  __shared__ uint32_t s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    s_last = atomicAdd(done, 1u) == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  const bool last = s_last != 0u;
  if (last) __threadfence();
  return last;
}

// Exclusive scan over both orders' counters, which lie one after the other
// (32 buckets per thread of a 1024-thread work-group, so thread 8 c starts
// code c of the join order; the join order counts every row, so the
// first-byte order's offsets are the scan less n), into dst (global memory
// or LDS); the start of each code's rows in the join order into
// s_code_start; with `clear`, the counters are left zero.  Ends with a
// barrier.
__device__ void lsmall_offsets(const LayoutSmall &a, const uint32_t *hist, uint32_t *dst,
                               bool clear, uint32_t *s_code_start) {
  __shared__ uint32_t s_part[32];  // wave totals, then their inclusive scan
  const uint32_t t = threadIdx.x;
  const uint32_t nb = a.n_codes * 256u;
  const uint32_t total = (a.sig ? 2u : 1u) * nb;  // a multiple of 256
  const uint32_t first = 32u * t;
  const bool mine = first < total;
  uint32_t v[32];
  if (mine) {
    const uint4 *h4 = reinterpret_cast<const uint4 *>(hist + first);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint4 x = h4[q];
      v[4 * q] = x.x;
      v[4 * q + 1] = x.y;
      v[4 * q + 2] = x.z;
      v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 32; q++) v[q] = 0;
  }
  uint32_t sum = 0;
#pragma unroll
  for (int q = 0; q < 32; q++) sum += v[q];
  // exclusive scan of the partial sums: within each wave by shuffles, then
  // over the 16 wave totals
  const uint32_t lane = lane_id(), wv = t >> 6;
  uint32_t inc = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)inc, d);
    if (lane >= (uint32_t)d) inc += y;
  }
  if (lane == 63) s_part[wv] = inc;
  __syncthreads();
  if (t < 64) {
    uint32_t w = t < 16 ? s_part[t] : 0u;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)w, d);
      if (t >= (uint32_t)d) w += y;
    }
    if (t < 16) s_part[16 + t] = w;  // inclusive over waves
  }
  __syncthreads();
  uint32_t run = inc - sum + (wv ? s_part[16 + wv - 1] : 0u) - (first >= nb ? a.n : 0u);
  if (first < nb && (t & 7u) == 0) s_code_start[t >> 3] = run;
  if (mine) {
    uint4 *o4 = reinterpret_cast<uint4 *>(dst + first);
    uint4 *z4 = reinterpret_cast<uint4 *>(const_cast<uint32_t *>(hist) + first);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      uint4 x;
      x.x = run;
      run += v[4 * q];
      x.y = run;
      run += v[4 * q + 1];
      x.z = run;
      run += v[4 * q + 2];
      x.w = run;
      run += v[4 * q + 3];
      o4[q] = x;
      if (clear) z4[q] = make_uint4(0u, 0u, 0u, 0u);
    }
  }
  __syncthreads();
}

// The runs of the join order from the code starts (one per code with rows;
// codes < 64: one wave's ballot) and their count, by the first wave.
__device__ void lsmall_runs(const LayoutSmall &a, const uint32_t *s_code_start) {
  const uint32_t t = threadIdx.x;
  if (t >= 64) return;
  const uint32_t c = t;
  const uint32_t lo = c < a.n_codes ? s_code_start[c] : a.n;
  const uint32_t hi = c + 1 < a.n_codes ? s_code_start[c + 1] : a.n;
  const bool has = c < a.n_codes && hi > lo;
  const uint64_t m = __ballot(has);
  const uint32_t id = rank_in(m);
  if (has) {
    a.run_code[id] = c;
    a.run_off[id] = lo;
  }
  if (t == 0) {
    const uint32_t nr = (uint32_t)__popcll(m);
    a.run_off[nr] = a.n;
    *a.n_runs = nr;
    *a.n_runs_host = nr;
  }
}

// Run by the global-atomic counting launch's last work-group (1024
// threads): the offsets into a.off, the runs, the layout's bad flag; the
// counters, the flag accumulator and the done count are left zero for the
// next layout built on the same stream.
__device__ void lsmall_scan(const LayoutSmall &a) {
  __shared__ uint32_t s_code_start[kLayoutMaxCodes + 1];
  lsmall_offsets(a, a.hist, a.off, true, s_code_start);
  lsmall_runs(a, s_code_start);
  if (threadIdx.x == 0) {
    *a.bad_out = *a.bad;
    *a.bad = 0u;
    *a.done = 0u;
  }
}

// Ranks from global atomics (keys of up to kLayoutMaxCodes codes).
__global__ __launch_bounds__(1024) void k_lsmall_count(LayoutSmall a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) {
    const uint32_t base = lsmall_code(a, i, a.bad) << 8;
    const uint32_t l = a.len ? min((uint32_t)a.len[i], 255u) : 0u;
    a.rank[i] = atomicAdd(a.hist + base + l, 1u);
    if (a.sig) {
      const uint32_t fb = a.sig[i].pre & 0xffu;
      a.rank[a.n + i] = atomicAdd(a.hist + a.n_codes * 256u + base + fb, 1u);
    }
  }
  if (last_group(a.done)) lsmall_scan(a);
}

// Layouts of up to kLayoutLdsCodes codes: the counts of a work-group's 1024
// rows (one per thread) taken in LDS first, one global atomic per bucket the
// work-group uses, into a.hist (zero on entry).  The launch also clears
// a.hist_clear (the counters of the previous such layout, which its
// k_lsmall_place has read) and the layout's bad flag; the offsets are
// scanned by every work-group of k_lsmall_place.
constexpr uint32_t kLayoutRowsPerGroup = 1024;
__global__ __launch_bounds__(1024) void k_lsmall_count_lds(LayoutSmall a) {
  extern __shared__ uint32_t s_cnt[];  // buckets of both orders
  const uint32_t t = threadIdx.x;
  const uint32_t nb = a.n_codes * 256u, nb2 = a.sig ? 2u * nb : nb;
  for (uint32_t k = t; k < nb2; k += 1024) s_cnt[k] = 0;
  for (uint32_t k = blockIdx.x * 1024 + t; k < a.clear_words; k += gridDim.x * 1024)
    a.hist_clear[k] = 0u;
  if (blockIdx.x == 0 && t == 0) *a.bad_out = 0u;  // k_lsmall_place sets it
  __syncthreads();
  const uint32_t i = blockIdx.x * kLayoutRowsPerGroup + t;
  uint32_t b1 = 0xffffffffu, b2 = 0xffffffffu, r1 = 0, r2 = 0;
  if (i < a.n) {
    const uint32_t base = lsmall_code(a, i, nullptr) << 8;
    const uint32_t l = a.len ? min((uint32_t)a.len[i], 255u) : 0u;
    b1 = base + l;
    if (a.sig) b2 = nb + base + (a.sig[i].pre & 0xffu);
  }
  if (b1 != 0xffffffffu) r1 = atomicAdd(s_cnt + b1, 1u);
  if (b2 != 0xffffffffu) r2 = atomicAdd(s_cnt + b2, 1u);
  __syncthreads();
  for (uint32_t k = t; k < nb2; k += 1024) {  // the work-group's base in each bucket
    const uint32_t c = s_cnt[k];
    s_cnt[k] = c ? atomicAdd(a.hist + k, c) : 0u;
  }
  __syncthreads();
  if (b1 != 0xffffffffu) a.rank[i] = s_cnt[b1] + r1;
  if (b2 != 0xffffffffu) a.rank[a.n + i] = s_cnt[b2] + r2;
}

// One row's places in both orders, its signature copied to each.
__device__ __forceinline__ void lsmall_put(const LayoutSmall &a, uint32_t i, const uint32_t *off) {
  const uint32_t base = lsmall_code(a, i, a.bad_out) << 8;
  const uint32_t l = a.len ? min((uint32_t)a.len[i], 255u) : 0u;
  const uint32_t pos = off[base + l] + a.rank[i];
  a.perm[pos] = i;
  if (a.sig) {
    // The row as three 16-byte words (word 8, the first 4 bytes, opens the
    // third), in clang vector types: a SigRow or uint4 temporary copied to
    // two places is given a per-thread copy in LDS by the compiler.
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 *src = reinterpret_cast<const u32x4 *>(a.sig + i);
    const u32x4 w0 = src[0], w1 = src[1], w2 = src[2];
    u32x4 *d = reinterpret_cast<u32x4 *>(a.jsig + pos);
    d[0] = w0;
    d[1] = w1;
    d[2] = w2;
    const uint32_t pos2 = off[a.n_codes * 256u + base + (w2.x & 0xffu)] + a.rank[a.n + i];
    a.perm_fb[pos2] = i;
    d = reinterpret_cast<u32x4 *>(a.jsig_fb + pos2);
    d[0] = w0;
    d[1] = w1;
    d[2] = w2;
  }
}

// Places the rows of a layout counted by k_lsmall_count_lds, one row per
// thread (the grid covers n): every work-group scans the counts (at most
// 2 * kLayoutLdsCodes * 256) into LDS itself, so no work-group waits for
// another, and the first also writes the runs.  A row's own reads are
// issued before the scan, so their latency overlaps it.
__global__ __launch_bounds__(1024) void k_lsmall_place(LayoutSmall a) {
  extern __shared__ uint32_t s_off[];
  __shared__ uint32_t s_code_start[kLayoutMaxCodes + 1];
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
  const bool has = i < a.n;
  uint32_t base = 0, l = 0, rk = 0, rk2 = 0;
  u32x4 w0 = {0u, 0u, 0u, 0u}, w1 = w0, w2 = w0;
  if (has) {
    base = lsmall_code(a, i, a.bad_out) << 8;
    l = a.len ? min((uint32_t)a.len[i], 255u) : 0u;
    rk = a.rank[i];
    if (a.sig) {
      rk2 = a.rank[a.n + i];
      const u32x4 *src = reinterpret_cast<const u32x4 *>(a.sig + i);
      w0 = src[0];
      w1 = src[1];
      w2 = src[2];
    }
  }
  lsmall_offsets(a, a.hist, s_off, false, s_code_start);
  if (blockIdx.x == 0) lsmall_runs(a, s_code_start);
  if (!has) return;
  const uint32_t pos = s_off[base + l] + rk;
  a.perm[pos] = i;
  if (a.sig) {
    u32x4 *d = reinterpret_cast<u32x4 *>(a.jsig + pos);
    d[0] = w0;
    d[1] = w1;
    d[2] = w2;
    const uint32_t pos2 = s_off[a.n_codes * 256u + base + (w2.x & 0xffu)] + rk2;
    a.perm_fb[pos2] = i;
    d = reinterpret_cast<u32x4 *>(a.jsig_fb + pos2);
    d[0] = w0;
    d[1] = w1;
    d[2] = w2;
  }
}

__global__ void k_lsmall_scatter(LayoutSmall a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) lsmall_put(a, i, a.off);
}

// Keys of the first-byte order: (code << 16) | (first byte << 8) |
// min(length, 255).  Sorted, a join group's rows come ordered by first byte.
__global__ void k_layout_keys_fb(const uint32_t *__restrict__ code, const SigRow *__restrict__ sig,
                                 const uint16_t *__restrict__ len, uint32_t n, uint64_t bound,
                                 uint32_t *bad, uint64_t *__restrict__ keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t c = checked_code(code, i, bound, bad);
  const uint32_t f = sig[i].pre & 0xffu;
  const uint32_t l = min((uint32_t)len[i], 255u);
  keys[i] = (c << 16) | (f << 8) | l;
}

__global__ void k_gather_sig(const uint32_t *__restrict__ perm, uint32_t n,
                             const SigRow *__restrict__ sig, SigRow *__restrict__ jsig) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) jsig[i] = sig[perm[i]];
}

__global__ void k_run_flags(const uint64_t *__restrict__ keys, uint32_t n, uint32_t *flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = (i == 0 || (keys[i] >> 8) != (keys[i - 1] >> 8)) ? 1u : 0u;
}

// First index in [lo, hi) whose code (key >> 8) exceeds `code`.
__device__ __forceinline__ uint32_t upper_bound_code(const uint64_t *a, uint32_t lo, uint32_t hi,
                                                     uint64_t code) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if ((a[mid] >> 8) <= code) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// First index in [lo, hi) whose code is not below `code`.
__device__ __forceinline__ uint32_t lower_bound_code(const uint64_t *a, uint32_t lo, uint32_t hi,
                                                     uint64_t code) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if ((a[mid] >> 8) < code) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void set_group_work(const GroupArrays &g, uint32_t gi, uint32_t l_cnt,
                                               uint32_t r_cnt) {
  const uint64_t pairs = (uint64_t)l_cnt * r_cnt;
  uint64_t small = 0, items = 0;
  if (pairs > g.small_pairs_max) {
    const uint32_t a = max(l_cnt, r_cnt), b = min(l_cnt, r_cnt);
    items = (uint64_t)((a + g.tile_a - 1) / g.tile_a) * ((b + g.tile_b - 1) / g.tile_b);
  } else {
    small = pairs;
  }
  g.small_pairs[gi] = small;
  g.items[gi] = items;
}

// Runs of equal code in a sorted layout: run g starts at run_off[g] with
// code run_code[g]; run_off[n_runs] = n.
__global__ void k_layout_runs(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ flags,
                              const uint32_t *__restrict__ run_id, uint32_t n,
                              uint32_t *__restrict__ run_code, uint32_t *__restrict__ run_off,
                              uint32_t *__restrict__ n_runs, uint32_t *n_runs_host) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags[i]) {
    run_code[run_id[i]] = (uint32_t)(keys[i] >> 8);
    run_off[run_id[i]] = i;
  }
  if (i == n - 1) {
    const uint32_t nr = run_id[i] + flags[i];
    run_off[nr] = n;
    *n_runs = nr;
    *n_runs_host = nr;
  }
}

// Largest j in [0, n) with off[j] <= x (off non-decreasing, off[0] <= x).
__device__ __forceinline__ uint32_t find_le_u32(const uint32_t *off, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (off[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// Largest j in [start, n) with off[j] <= x, given off[start] <= x: steps of
// 1, 2, 4, ... from start, then a binary search of the last step.
__device__ __forceinline__ uint32_t find_le_from(const uint32_t *off, uint32_t n, uint32_t x,
                                                 uint32_t start) {
  uint32_t lo = start, step = 1;
  while (lo + step < n && off[lo + step] <= x) {
    lo += step;
    step <<= 1;
  }
  uint32_t hi = min(n, lo + step);
  while (hi - lo > 1) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (off[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// One thread per group of the batch: left run `local` of JR jr meets the
// right run of the same code (binary search over the right runs).
__device__ __forceinline__ void join_group(const BatchDev &b, const GroupArrays &g, uint32_t gi,
                                           uint32_t &job, unsigned long long &jp) {
  if (gi == g.n) {
    g.l_off[gi] = g.l_cnt[gi] = g.r_off[gi] = g.r_cnt[gi] = g.jr[gi] = 0;
    g.small_pairs[gi] = 0;
    g.items[gi] = 0;
    return;
  }
  // The JR of the wave's first active group by a binary search, each lane's
  // from there: a wave's groups are consecutive, in one JR or a few.
  const uint32_t g_first = __builtin_amdgcn_readfirstlane(gi);
  const uint32_t jr0 = find_le_u32(b.jr_group_off, b.n_jr, g_first);
  const uint32_t jr = find_le_from(b.jr_group_off, b.n_jr, gi, jr0);
  const JRDesc &d = b.jrs[jr];
  const uint32_t local = gi - b.jr_group_off[jr];
  const SideDesc &L = d.s[0], &R = d.s[1];
  if (local == 0 && (*L.bad | *R.bad)) atomicAdd(b.invalid, 1ull);
  if (local >= *L.n_runs_dev) {  // the host sized this JR's groups by a bound
    g.l_off[gi] = g.l_cnt[gi] = g.r_off[gi] = g.r_cnt[gi] = 0;
    g.jr[gi] = jr;
    g.small_pairs[gi] = 0;
    g.items[gi] = 0;
    return;
  }
  const uint32_t code = L.run_code[local];
  const uint32_t l_off = L.run_off[local], l_cnt = L.run_off[local + 1] - l_off;
  const uint32_t r_runs = *R.n_runs_dev;
  uint32_t lo = 0, hi = r_runs;
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (R.run_code[mid] < code) lo = mid + 1; else hi = mid;
  }
  uint32_t r_off = 0, r_cnt = 0;
  if (lo < r_runs && R.run_code[lo] == code) {
    r_off = R.run_off[lo];
    r_cnt = R.run_off[lo + 1] - r_off;
  }
  g.l_off[gi] = l_off;
  g.l_cnt[gi] = l_cnt;
  g.r_off[gi] = r_off;
  g.r_cnt[gi] = r_cnt;
  g.jr[gi] = jr;
  set_group_work(g, gi, l_cnt, r_cnt);
  job = d.job;
  jp = (unsigned long long)l_cnt * r_cnt;
}

// One thread per group (and one more for the end entry).  The candidate
// pairs are summed per job over the wave, one atomic per job and wave.
__global__ void k_batch_join(const BatchDev b, const GroupArrays g) {
  const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t job = 0;
  unsigned long long jp = 0;
  if (gi <= g.n) join_group(b, g, gi, job, jp);
  job_add(b, kJobJoined, jp != 0, job, jp);
}

// A batch whose groups (and end entry) fit one work-group of 1024 starts
// with this one launch in place of four: its descriptors copied from the
// pinned host buffer (read over the bus, as the copy would read them), its
// counters and dedupe sets cleared, the join (one thread per group, as
// k_batch_join) and the exclusive scan of the groups' work items.  The
// work-group's barriers order each phase's global writes before the next
// phase's reads.
__global__ __launch_bounds__(1024) void k_batch_prologue(const BatchDev b, const GroupArrays g,
                                                        const uint4 *__restrict__ hdesc,
                                                        uint4 *__restrict__ ddesc, uint32_t desc16,
                                                        unsigned long long *__restrict__ cnt,
                                                        uint32_t cnt_words, uint64_t set_words) {
  __shared__ unsigned long long s_wave[16];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < desc16; i += 1024) ddesc[i] = hdesc[i];
  for (uint32_t i = t; i < cnt_words; i += 1024) cnt[i] = 0ull;
  for (uint64_t i = t; i < set_words; i += 1024) b.sets.slots[i] = ~0ull;
  __syncthreads();
  uint32_t job = 0;
  unsigned long long jp = 0;
  if (t <= g.n) join_group(b, g, t, job, jp);
  job_add(b, kJobJoined, jp != 0, job, jp);
  __syncthreads();
  const uint32_t lane = lane_id(), wave = t >> 6;
  const unsigned long long v = t <= g.n ? g.items[t] : 0ull;
  unsigned long long inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(inc, d);
    if (lane >= (uint32_t)d) inc += y;
  }
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  unsigned long long base = 0;
  for (uint32_t w = 0; w < wave; w++) base += s_wave[w];
  if (t <= g.n) g.item_off[t] = base + inc - v;
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

// ---- stage 1: bound filter over join groups ---------------------------------

// Signatures arrive in join order (the JR's layouts): element i of jsig
// belongs to row perm[i], so the rows of a join group are contiguous.

// The bound test of may_pass in allowance form (hb_allow.h): for a
// lane-held A row and a B row shared by the wave, the pair is rejected iff
// SAD > allow(la, lb), found by the same float comparison, so decisions
// equal the direct form bit for bit.

// Negated allowances of one A row for one B length, both prefix classes.
__device__ __forceinline__ void allowances(uint32_t la, float rla, uint32_t lb, float rlb,
                                           float k0, float k3, uint32_t &nal0, uint32_t &nal3) {
  const float s = rla + rlb;
  const float inv = s < INFINITY ? __builtin_amdgcn_rcpf(s) : 0.0f;
  allowances_inv(la, rla, lb, rlb, k0, k3, inv, nal0, nal3);
}

// The allowances of every (A length, B length) pair for one threshold, so
// that the filter reads them per run instead of computing them.  Entry
// (lb, la) is what `allowances` gives for signatures of those lengths (the
// reciprocal formed as in k_build_sig), so the table changes no decision.
__global__ __launch_bounds__(256) void k_allow_table(float c0, float c3, uint32_t *__restrict__ tab) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kAllowRows * kAllowCols) return;
  const uint32_t lb = i / kAllowCols, la = i % kAllowCols;
  uint32_t n0 = kRejectAll, n3 = kRejectAll;
  if (la < 256) {
    const float rla = (la >= 1u && la <= 255u) ? 1.0f / (float)la : INFINITY;
    const float rlb = (lb >= 1u && lb <= 255u) ? 1.0f / (float)lb : INFINITY;
    allowances(la, rla, lb, rlb, 2.0f * c0, 2.0f * c3, n0, n3);
  }
  tab[2 * i] = n0;
  tab[2 * i + 1] = n3;
}

// The accumulator chain: seed + sum of |a_i - b_i| over the 32 counters.
__device__ __forceinline__ int32_t sad_chain(const uint32_t (&ah)[8], const uint32_t (&bh)[8],
                                             uint32_t seed) {
  uint32_t acc = seed;
#pragma unroll
  for (int k = 0; k < 8; k++) acc = __builtin_amdgcn_sad_u8(ah[k], bh[k], acc);
  return (int32_t)acc;
}

// coef[0] applies when the first bytes differ, coef[3] (the smallest of
// coef[1..3], so never stricter) when only a later one of the first 4 bytes
// differs, and nothing is rejected when the first 4 bytes are equal.  With
// the prefixes byte-swapped (first byte most significant), x = pre_a ^ pre_b
// is >= 2^24 exactly when the first bytes differ and 0 when all 4 are equal.
// Returns the accumulator: the pair is rejected iff it is > 0 and x != 0.
__device__ __forceinline__ int32_t bound_acc(const uint32_t (&ah)[8], uint32_t nal0, uint32_t nal3,
                                             const uint32_t (&bh)[8], uint32_t x) {
  uint32_t acc = x >= 0x01000000u ? nal0 : nal3;
#pragma unroll
  for (int k = 0; k < 8; k++) acc = __builtin_amdgcn_sad_u8(ah[k], bh[k], acc);
  return (int32_t)acc;
}

// Work item = 256*AR rows of the larger side in lanes (A, AR rows per lane)
// x up to tile_b rows of the other side (B).  B rows are staged through LDS
// in chunks of 256 and read as wave-wide broadcasts, two at a time, so each
// LDS round trip feeds 2*AR independent pair tests.  Inside a join group the
// rows are ordered by length, so a chunk splits into a few runs of equal B
// length; the per-lane allowances are computed once per run and the rows of
// a run are tested without further branching.  Every pair of a group
// already satisfies all of its rule's equalities (they form the join key).
// B row as staged in LDS: byte-class counters, byte-swapped prefix, length.
struct __attribute__((aligned(8))) BRow {
  uint32_t h[8];
  uint32_t pre;
  uint32_t len;
};

// Read-only data the kernel reads at wave-uniform addresses: through the
// constant address space such loads become scalar loads.
typedef const __attribute__((address_space(4))) SigRow ConstSigRow;

// Small groups (at most small_pairs_max pairs each), run by the filter
// launch's last work-groups (k_filter_tiles): one wave per group, groups
// taken in a grid-stride loop over those waves (wbuf: the wave's output
// buffer of kWaveOut entries in LDS).  The larger side sits in lanes, 64
// rows at a time; the other side's rows are read with scalar loads in join
// order, where they are ordered by length, so the lanes' allowances are
// recomputed only where the length changes.
template <bool SIM>
__device__ void filter_small_groups(const BatchDev &bd, const GroupArrays &g, uint32_t g_begin,
                                    uint32_t g_end, const PairList &out, PairEntry *wbuf,
                                    uint32_t wave_in, uint32_t n_waves) {
  uint32_t wn = 0;  // entries in wbuf (wave-uniform)
  const uint32_t lane = lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(wave_in);
  for (uint32_t gi = g_begin + wave; gi < g_end; gi += n_waves) {
    if (__builtin_amdgcn_readfirstlane((uint32_t)g.small_pairs[gi]) == 0 &&
        __builtin_amdgcn_readfirstlane((uint32_t)(g.small_pairs[gi] >> 32)) == 0)
      continue;  // tiled or empty
    const uint32_t jr = __builtin_amdgcn_readfirstlane(g.jr[gi]);
    const uint32_t l_off = __builtin_amdgcn_readfirstlane(g.l_off[gi]);
    const uint32_t l_cnt = __builtin_amdgcn_readfirstlane(g.l_cnt[gi]);
    const uint32_t r_off = __builtin_amdgcn_readfirstlane(g.r_off[gi]);
    const uint32_t r_cnt = __builtin_amdgcn_readfirstlane(g.r_cnt[gi]);
    const JRDesc *d = bd.jrs + jr;
    const bool a_is_l = l_cnt >= r_cnt;
    const uint32_t a_cnt = a_is_l ? l_cnt : r_cnt, b_cnt = a_is_l ? r_cnt : l_cnt;
    const uint32_t a_off = a_is_l ? l_off : r_off, b_off = a_is_l ? r_off : l_off;
    const SideDesc &sa = d->s[a_is_l ? 0 : 1], &sb = d->s[a_is_l ? 1 : 0];
    const uint32_t *perm_a = wave_ptr(sa.perm) + a_off;
    const uint32_t *perm_b = wave_ptr(sb.perm) + b_off;
    const SigRow *sig_a = SIM ? wave_ptr(sa.jsig) + a_off : nullptr;
    ConstSigRow *sig_b = SIM ? (ConstSigRow *)(wave_ptr(sb.jsig) + b_off) : nullptr;
    const float k0 = 2.0f * d->c0, k3 = 2.0f * d->c3;
    for (uint32_t a0 = 0; a0 < a_cnt; a0 += 64) {
      const uint32_t ai = a0 + lane;
      const bool av = ai < a_cnt;
      uint32_t ah[8] = {0, 0, 0, 0, 0, 0, 0, 0}, apre = 0, la = 0;
      float rla = INFINITY;
      if (SIM && av) {
        const SigRow s = sig_a[ai];
#pragma unroll
        for (int k = 0; k < 8; k++) ah[k] = s.h[k];
        apre = __builtin_bswap32(s.pre);
        la = s.len;
        rla = s.rlen;
      }
      uint32_t cur_len = 0xffffffffu, nal0 = (uint32_t)(-kPassAllow), nal3 = nal0;
      for (uint32_t bj = 0; bj < b_cnt; bj++) {
        bool ok = av;
        if (SIM) {
          uint32_t bh[8];
#pragma unroll
          for (int k = 0; k < 8; k++) bh[k] = sig_b[bj].h[k];
          const uint32_t blen = sig_b[bj].len;
          if (blen != cur_len) {
            cur_len = blen;
            allowances(la, rla, blen, sig_b[bj].rlen, k0, k3, nal0, nal3);
          }
          const uint32_t x = apre ^ __builtin_bswap32(sig_b[bj].pre);
          ok = ok && (bound_acc(ah, nal0, nal3, bh, x) <= 0 || x == 0u);
        }
        if (__builtin_amdgcn_ballot_w64(ok)) {
          const uint32_t b_row = perm_b[bj];
          const uint32_t a_row = ok ? perm_a[ai] : 0u;
          wave_append(ok, PairEntry{a_is_l ? a_row : b_row, a_is_l ? b_row : a_row, jr, 0u},
                      wbuf, wn, out, bd.overflow);
        }
      }
    }
  }
  wave_flush(wbuf, wn, out, bd.overflow);
}

template <bool SIM, bool TAB, int AR, int BU>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AR == 4 ? 5 : (AR == 2 ? 7 : 8)))) void
k_filter_tiles(const BatchDev bd, const GroupArrays g, uint32_t g_begin, uint32_t g_end,
               PairList out, uint32_t tile_grid) {
  constexpr uint32_t kChunk = 256;
  __shared__ BRow s_b[SIM ? kChunk + 1 : 1];
  __shared__ unsigned long long s_starts[kChunk / 64];  // bit i: B row i starts a run
  __shared__ float s_rlen[SIM ? kChunk + 1 : 1];
  // First-byte range of each slice (64 A rows: wave w, row r of its lanes),
  // and per slice a bit per B row of the chunk whose first byte lies in it.
  constexpr int kSlices = SIM ? 4 * AR : 1;
  __shared__ uint32_t s_frange[kSlices];
  __shared__ unsigned long long s_slow[kSlices][kChunk / 64];
  __shared__ TileOut to;
  const uint32_t tid = threadIdx.x;
  if (blockIdx.x >= tile_grid) {  // the launch's last work-groups: small groups
    static_assert(sizeof(TileOut::e) >= 4 * kWaveOut * sizeof(PairEntry), "small-path buffer");
    filter_small_groups<SIM>(bd, g, g_begin, g_end, out,
                             reinterpret_cast<PairEntry *>(to.e) + (tid >> 6) * kWaveOut,
                             (blockIdx.x - tile_grid) * 4u + (tid >> 6),
                             (gridDim.x - tile_grid) * 4u);
    return;
  }
  if (tid == 0) {
    to.n = 0;
    to.dense = 0;
  }
  const uint32_t tile_b = g.tile_b;
  // Item-level values are wave-uniform; they are moved to scalar registers
  // so that the per-pair loop keeps its vector registers.
  const uint64_t item_begin = wave_u64(g.item_off[g_begin]);
  const uint64_t item_end = wave_u64(g.item_off[g_end]);
  for (uint64_t item = item_begin + blockIdx.x; item < item_end; item += tile_grid) {
    const uint32_t gi =
        wave_u32(g_begin + find_group(g.item_off + g_begin, g_end - g_begin, item));
    const uint32_t jr = wave_u32(g.jr[gi]);
    const JRDesc &d = bd.jrs[jr];
    const float k0 = wave_f32(2.0f * d.c0), k3 = wave_f32(2.0f * d.c3);
    const uint32_t *atab = TAB ? wave_ptr(d.allow_tab) : nullptr;
    const bool dense_item = !SIM || wave_u32(d.refine) != 0;
    const uint32_t l_off = wave_u32(g.l_off[gi]), l_cnt = wave_u32(g.l_cnt[gi]);
    const uint32_t r_off = wave_u32(g.r_off[gi]), r_cnt = wave_u32(g.r_cnt[gi]);
    const bool a_is_l = l_cnt >= r_cnt;
    const uint32_t a_cnt = a_is_l ? l_cnt : r_cnt, b_cnt = a_is_l ? r_cnt : l_cnt;
    const uint32_t a_off = a_is_l ? l_off : r_off, b_off = a_is_l ? r_off : l_off;
    const SideDesc &sa = d.s[a_is_l ? 0 : 1], &sb = d.s[a_is_l ? 1 : 0];
    // A rows in first-byte order, B rows in length order (runs of equal length)
    const uint32_t *perm_a = wave_ptr(sa.perm_fb);
    const uint32_t *perm_b = wave_ptr(sb.perm);
    const SigRow *sig_a = wave_ptr(sa.jsig_fb) + a_off;
    const SigRow *sig_b = wave_ptr(sb.jsig) + b_off;
    const uint64_t local = item - wave_u64(g.item_off[gi]);
    const uint32_t n_bseg = (b_cnt + tile_b - 1) / tile_b;
    const uint32_t a_tile = (uint32_t)(local / n_bseg), b_seg = (uint32_t)(local % n_bseg);
    // Rows of the A tile by wave: wave w holds rows [64*AR*w, 64*AR*(w+1)),
    // so a partly filled last tile leaves whole waves idle, and idle waves
    // skip the pair tests (they still take part in the group's barriers).
    const uint32_t wave = wave_u32(tid >> 6);
    const uint32_t a_wave0 = a_tile * 256 * AR + wave * 64 * AR;
    const bool wave_active = a_wave0 < a_cnt;  // wave-uniform (scalar operands)
    bool a_valid[AR];
    uint32_t a_idx[AR], ah[AR][8], apre[AR], la[AR];
    uint32_t nal0[AR], nal3[AR];
    float rla[AR];
#pragma unroll
    for (int r = 0; r < AR; r++) {
      a_idx[r] = a_wave0 + r * 64 + lane_id();
      a_valid[r] = a_idx[r] < a_cnt;
#pragma unroll
      for (int k = 0; k < 8; k++) ah[r][k] = 0;
      apre[r] = 0xffffffffu;
      la[r] = 0;
      rla[r] = INFINITY;
      nal0[r] = nal3[r] = kRejectAll;
      if (SIM && a_valid[r]) {
        const SigRow s = sig_a[a_idx[r]];
#pragma unroll
        for (int k = 0; k < 8; k++) ah[r][k] = s.h[k];
        apre[r] = __builtin_bswap32(s.pre);
        la[r] = s.len & 0xffffu;
        rla[r] = s.rlen;
      }
      if (TAB) {  // bits 16 and up: byte offset of the row's allowance-table column
        const uint32_t col = !a_valid[r] ? 256u : (rla[r] < INFINITY ? la[r] : 0u);
        la[r] |= (col * 8u) << 16;
      }
    }
    if (SIM) {
      // The slice's valid rows are consecutive in first-byte order, so their
      // first bytes span [first lane's, last valid lane's]; an empty slice
      // gets an empty range.  (The previous item's last reads of s_frange
      // precede the barrier that ended it.)
#pragma unroll
      for (int r = 0; r < AR; r++) {
        const uint32_t base = a_wave0 + r * 64;
        const uint32_t nv = base < a_cnt ? min(64u, a_cnt - base) : 0u;
        const uint32_t fb = apre[r] >> 24;
        const uint32_t lo = nv ? (uint32_t)__builtin_amdgcn_readlane((int)fb, 0) : 256u;
        const uint32_t hi = nv ? (uint32_t)__builtin_amdgcn_readlane((int)fb, (int)nv - 1) : 0u;
        if (lane_id() == 0) s_frange[wave * AR + r] = lo | (hi << 16);
      }
    }
    const uint32_t b_begin = b_seg * tile_b;
    const uint32_t b_end = min(b_begin + tile_b, b_cnt);
    const TileCtx tc{perm_a + a_off, perm_b + b_off + b_begin, a_tile * 256 * AR, jr, a_is_l};
    for (uint32_t cb = b_begin; cb < b_end; cb += kChunk) {
      const uint32_t n_b = min(kChunk, b_end - cb);
      __syncthreads();
      if (to.n >= kTileOut / 2) tile_flush(to, tc, out, bd.overflow);  // to.n is group-uniform here
      uint32_t my_len = 0, my_fb = 0;
      if (tid < n_b) {
        if (SIM) {
          const SigRow bs = sig_b[cb + tid];
          BRow br;
#pragma unroll
          for (int k = 0; k < 8; k++) br.h[k] = bs.h[k];
          br.pre = __builtin_bswap32(bs.pre);
          br.len = my_len = bs.len;
          my_fb = bs.pre & 0xffu;
          s_b[tid] = br;
          s_rlen[tid] = bs.rlen;
        }
      }
      __syncthreads();
      if (SIM) {  // run starts: rows whose length differs from the previous row's
        const bool start = tid < n_b && (tid == 0 || s_b[tid - 1].len != my_len);
        const unsigned long long m = __ballot(start);
        if (lane_id() == 0) s_starts[tid >> 6] = m;
#pragma unroll
        for (int sl = 0; sl < kSlices; sl++) {
          const uint32_t fr = s_frange[sl];
          const bool slow = tid < n_b && my_fb >= (fr & 0xffffu) && my_fb <= (fr >> 16);
          const unsigned long long ms = __ballot(slow);
          if (lane_id() == 0) s_slow[sl][tid >> 6] = ms;
        }
        __syncthreads();
      }
      // read once per chunk: `to.dense` changes only inside flushes, and a
      // chunk that starts sparse flushes only at its end
      const bool dense = dense_item || wave_u32(to.dense) != 0;
      // Without flushes between steps an idle wave has nothing to do here.
      if (!dense && !wave_active) continue;
      // One step: the lanes' A rows against B rows jj .. jj+NB-1 of the
      // chunk.  A pair is kept when its accumulator is <= 0.  Where a B
      // row's first byte lies outside the first-byte range of a slice (bit
      // clear in `slow`), every pair of the two differs in the first byte:
      // the accumulator starts from nal0 with no per-pair class test.
      // Otherwise the class is chosen per pair and a pair whose first 4
      // bytes are equal is kept.  The step's common case (no pair kept) is
      // told by the minimum over the step.
      auto step = [&](uint32_t jj, auto nb_tag, const uint64_t (&slow_w)[AR]) {
        constexpr int NB = decltype(nb_tag)::value;
        int32_t acc[NB][AR];
        uint64_t slow_any = 0;
#pragma unroll
        for (int r = 0; r < AR; r++) slow_any |= (slow_w[r] >> (jj & 63u)) & ((1ull << NB) - 1);
        if (!SIM) {
#pragma unroll
          for (int u = 0; u < NB; u++)
#pragma unroll
            for (int r = 0; r < AR; r++) acc[u][r] = 0;
        } else if (slow_any == 0) {
#pragma unroll
          for (int u = 0; u < NB; u++) {
            const BRow b = s_b[jj + u];
#pragma unroll
            for (int r = 0; r < AR; r++) acc[u][r] = sad_chain(ah[r], b.h, nal0[r]);
          }
        } else {
#pragma unroll
          for (int u = 0; u < NB; u++) {
            const BRow b = s_b[jj + u];
#pragma unroll
            for (int r = 0; r < AR; r++) {
              const uint32_t x = apre[r] ^ b.pre;
              const int32_t a = sad_chain(ah[r], b.h, x >= 0x01000000u ? nal0[r] : nal3[r]);
              acc[u][r] = x == 0u ? 0 : a;
            }
          }
        }
        int32_t acc_min = acc[0][0];
#pragma unroll
        for (int u = 0; u < NB; u++)
#pragma unroll
          for (int r = 0; r < AR; r++) acc_min = min(acc_min, acc[u][r]);
        if (__builtin_amdgcn_ballot_w64(acc_min <= 0)) {  // rare: a survivor
#pragma unroll
          for (int u = 0; u < NB; u++)
#pragma unroll
            for (int r = 0; r < AR; r++)
              tile_append(a_valid[r] && acc[u][r] <= 0, a_idx[r] - tc.a_base,
                          cb - b_begin + jj + u, tc, to, out, bd.overflow);
        }
      };
      uint32_t run = 0;
      while (run < n_b) {  // wave-uniform
        uint32_t run_end = n_b;
        if (SIM) {
          if (wave_active) {
            const uint32_t blen = wave_u32(s_b[run].len);
            if (TAB) {
              const uint32_t bcol = blen - 1u < 255u ? blen : 0u;
              const char *row = (const char *)(atab + 2u * kAllowCols * bcol);
#pragma unroll
              for (int r = 0; r < AR; r++) {
                const uint2 e = *(const uint2 *)(row + (la[r] >> 16));
                nal0[r] = e.x;
                nal3[r] = e.y;
              }
            } else {
              const float rlb = s_rlen[run];
#pragma unroll
              for (int r = 0; r < AR; r++) {
                allowances(la[r], rla[r], blen, rlb, k0, k3, nal0[r], nal3[r]);
                if (!a_valid[r]) nal0[r] = nal3[r] = kRejectAll;
              }
            }
          }
          // next set bit of s_starts after `run`
          for (uint32_t w = (run + 1) >> 6; w < kChunk / 64; w++) {
            unsigned long long m = wave_u64(s_starts[w]);
            if (w == (run + 1) >> 6) m &= ~0ull << ((run + 1) & 63);
            if (m) {
              run_end = min(n_b, (w << 6) + (uint32_t)__builtin_ctzll(m));
              break;
            }
          }
        }
        // The run in segments within one 64-row word of the slow masks.
        for (uint32_t jj = run; jj < run_end;) {
          const uint32_t seg_end = min(run_end, (jj & ~63u) + 64u);
          uint64_t slow_w[AR];
#pragma unroll
          for (int r = 0; r < AR; r++)
            slow_w[r] = SIM ? wave_u64(s_slow[SIM ? wave * AR + r : 0][jj >> 6]) : 0ull;
          // Where survivors are dense (no similarity test, a loose first
          // threshold, or a large last flush) the buffer is also flushed
          // between steps: a step adds at most 256 * AR * BU <= kTileOut / 2
          // entries, so it never overflows.  Every wave of the group runs
          // the same (run, jj) sequence and reads the same `dense` (written
          // only between barriers), so all reach the barrier together.  The
          // count is read between two barriers: a wave that goes on to its
          // next step appends (changing to.n) only after every wave has
          // read it, so all waves take the same flush decision.
          auto between = [&](uint32_t j) {
            if (dense && j != run) {
              __syncthreads();
              const bool flush = to.n >= kTileOut / 2;
              __syncthreads();
              if (flush) tile_flush(to, tc, out, bd.overflow);
            }
          };
          if (!dense) {  // only active waves get here
            for (; jj + BU <= seg_end; jj += BU) step(jj, std::integral_constant<int, BU>{}, slow_w);
            if (BU > 1 && jj < seg_end) {  // one B row left in the segment
              step(jj, std::integral_constant<int, 1>{}, slow_w);
              jj++;
            }
          } else {
            for (; jj + BU <= seg_end; jj += BU) {
              between(jj);
              if (wave_active) step(jj, std::integral_constant<int, BU>{}, slow_w);
            }
            if (BU > 1 && jj < seg_end) {
              between(jj);
              if (wave_active) step(jj, std::integral_constant<int, 1>{}, slow_w);
              jj++;
            }
          }
        }
        run = run_end;
      }
    }
    __syncthreads();
    if (to.n) tile_flush(to, tc, out, bd.overflow);  // group-uniform after the barrier
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
  BatchDev b;
  PairList in;
  int stage;        // 1-based index of the similarity predicate evaluated
  int count_filtered;  // count the input as the filter's survivors (no refine pass ran)
  PairList next;    // pairs with further predicates
  PairList final_;  // pairs that passed their JR's last predicate (tagged with the job)
  PairList defer;   // pairs whose shorter string exceeds 256 bytes
  AuditList audit;
  // A batch of one job: final pairs and audit entries are written here, at
  // the positions they take in final_ and audit (the same capacities).
  uint2 *direct_pairs;
  hb_audit_entry *direct_audit;
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Inserts (l, r) into the job's output set; true if it was not there yet.
// Jobs with one rule have no set (each pair is evaluated once) and always
// insert.  A full set raises the overflow flag.
__device__ bool set_insert(const BatchDev &b, uint32_t job, uint32_t l, uint32_t r) {
  const unsigned long long cap = b.sets.cap ? b.sets.cap[job] : 0ull;
  if (cap == 0) return true;
  unsigned long long *slots = b.sets.slots + b.sets.off[job];
  const unsigned long long key = ((unsigned long long)l << 32) | r;
  unsigned long long slot = mix64(key) & (cap - 1);
  // Sets are sized for a load of at most one half: a probe sequence of 512
  // slots means the set is (nearly) full, and the batch reruns with a larger
  // one.
  const unsigned long long max_probe = cap < 512 ? cap : 512;
  for (unsigned long long probe = 0; probe < max_probe; probe++) {
    const unsigned long long prev = atomicCAS(slots + slot, kEmpty, key);
    if (prev == kEmpty) return true;
    if (prev == key) return false;
    slot = (slot + 1) & (cap - 1);
  }
  *b.overflow = 1ull;
  *b.set_overflow = 1ull;
  return false;
}

// A pair that satisfies a whole rule: appended to the final list once per
// job (called by every lane of the wave).
__device__ __forceinline__ void finalize(const ExactArgs &a, bool ok, uint32_t l, uint32_t r,
                                         uint32_t job) {
  const bool emit = ok && set_insert(a.b, job, l, r);
  if (!a.direct_pairs) {
    list_append(emit, PairEntry{l, r, job, 0u}, a.final_, a.b.overflow);
  } else {
    const uint64_t mask = __ballot(emit);
    if (mask) {
      const int leader = __ffsll((unsigned long long)mask) - 1;
      unsigned long long base = 0;
      if ((int)lane_id() == leader)
        base = atomicAdd(a.final_.count, (unsigned long long)__popcll(mask));
      base = readlane64(base, leader);
      if (emit) {
        const unsigned long long pos = base + rank_in(mask);
        if (pos < a.final_.cap)
          a.direct_pairs[pos] = make_uint2(l, r);
        else
          *a.b.overflow = 1ull;
      }
    }
  }
  job_count(a.b, kJobMatches, emit, job);
}

// Decision and audit entry of one evaluated pair (lane-local values, called
// by every lane of the wave); true when the pair passes the predicate.
__device__ __forceinline__ bool decide(const ExactArgs &a, const PairEntry &e, uint32_t job,
                                       float thr_f, int32_t slot, int m, int S, int T, int p,
                                       bool ok) {
  const double thr = (double)thr_f;
  const float jaro = jaro_value(m, S, T);
  const double jw = winkler(jaro, p);
  const bool pass = ok && !(jw < thr);
  const bool band = ok && fabs(jw - thr) < 1e-6;
  if (band) {
    const double jf = winkler_fused(jaro, p);
    const unsigned long long pos = atomicAdd(a.audit.count, 1ull);
    if (pos < a.audit.cap) {
      hb_audit_entry ae;
      ae.l = e.l;
      ae.r = e.r;
      ae.slot = slot;
      ae.flags = (((jw < thr) != (jf < thr)) ? 1 : 0) | (S >= 257 ? 2 : 0);
      ae.jw = jw;
      if (a.direct_audit) {
        a.direct_audit[pos] = ae;
      } else {
        AuditRec &rec = a.audit.e[pos];
        rec.e = ae;
        rec.job = job;
        rec.pad = 0;
      }
    } else {
      *a.b.overflow = 1ull;
    }
  }
  job_count(a.b, kJobAudit, band, job);
  return pass;
}

// Output of one work-group pass of k_exact_two (up to 256 pairs): entries for
// the next stage and for the final list are gathered in LDS and leave with
// one global atomic per list; the final entries' match count is added to
// their job with one atomic when they all belong to one job.
constexpr uint32_t kNoJob = 0xfffffffeu, kMixedJobs = 0xffffffffu;
struct StageOut {
  PairEntry next[256], fin[256];
  uint32_t n_next, n_fin;
  uint32_t fin_job;  // kNoJob, the one job of the final entries, or kMixedJobs
  unsigned long long base_next, base_fin;
};

__device__ __forceinline__ void stage_put(bool ok, const PairEntry &e, PairEntry *buf,
                                          uint32_t &n) {
  const uint64_t mask = __ballot(ok);
  if (mask == 0) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = atomicAdd(&n, (uint32_t)__popcll(mask));
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
  if (ok) buf[base + rank_in(mask)] = e;
}

// finalize() for k_exact_two: into the group's final buffer.
__device__ __forceinline__ void stage_final(const ExactArgs &a, StageOut &so, bool ok, uint32_t l,
                                            uint32_t r, uint32_t job) {
  const bool emit = ok && set_insert(a.b, job, l, r);
  const uint64_t mask = __ballot(emit);
  if (mask == 0) return;
  const int leader = __ffsll((unsigned long long)mask) - 1;
  const uint32_t j0 = (uint32_t)__builtin_amdgcn_readlane((int)job, leader);
  const uint32_t wave_job = __ballot(emit && job != j0) ? kMixedJobs : j0;
  if ((int)lane_id() == leader) {
    const uint32_t old = atomicCAS(&so.fin_job, kNoJob, wave_job);
    if (old != kNoJob && old != wave_job) so.fin_job = kMixedJobs;
  }
  stage_put(emit, PairEntry{l, r, job, 0u}, so.fin, so.n_fin);
}

// Moves the group's buffers to the global lists (called by every thread).
__device__ __forceinline__ void stage_flush(const ExactArgs &a, StageOut &so) {
  __syncthreads();
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    so.base_next = so.n_next ? atomicAdd(a.next.count, (unsigned long long)so.n_next) : 0ull;
    so.base_fin = so.n_fin ? atomicAdd(a.final_.count, (unsigned long long)so.n_fin) : 0ull;
    if (so.n_fin && so.fin_job != kMixedJobs)
      atomicAdd(a.b.job_cnt + (size_t)kJobMatches * a.b.cnt_stride + so.fin_job,
                (unsigned long long)so.n_fin);
  }
  __syncthreads();
  if (tid < so.n_next) {
    const unsigned long long pos = so.base_next + tid;
    if (pos < a.next.cap)
      a.next.e[pos] = so.next[tid];
    else
      *a.b.overflow = 1ull;
  }
  if (tid < so.n_fin) {
    const PairEntry e = so.fin[tid];
    const unsigned long long pos = so.base_fin + tid;
    if (pos >= a.final_.cap)
      *a.b.overflow = 1ull;
    else if (a.direct_pairs)
      a.direct_pairs[pos] = make_uint2(e.l, e.r);
    else
      a.final_.e[pos] = e;
    if (so.fin_job == kMixedJobs)  // the final entry's jr field holds its job
      atomicAdd(a.b.job_cnt + (size_t)kJobMatches * a.b.cnt_stride + e.jr, 1ull);
  }
  __syncthreads();
}

// Common prefix of a and b, capped at n and at 16 bytes (the cells' first
// 16 bytes, which every cell has).  The cap changes no result: from p = 11
// on, j + (1 - j) * p * 0.1 exceeds 1 for every Jaro value j < 1 (by
// 0.1 * (1 - j) >= 0.1 * 2^-24, far above the rounding) and equals 1 at
// j = 1, so Winkler's clamp yields exactly 1.0 either way.
__device__ __forceinline__ int lane_prefix(const uint8_t *a, const uint8_t *b, int n) {
  const uint4 x = *reinterpret_cast<const uint4 *>(a), y = *reinterpret_cast<const uint4 *>(b);
  const uint64_t lo = ((uint64_t)(x.y ^ y.y) << 32) | (uint64_t)(x.x ^ y.x);
  const uint64_t hi = ((uint64_t)(x.w ^ y.w) << 32) | (uint64_t)(x.z ^ y.z);
  const int p = lo ? (__builtin_ctzll(lo) >> 3) : (hi ? 8 + (__builtin_ctzll(hi) >> 3) : 16);
  return min(p, n);
}

// Removes the lowest set bit of the NW-word concatenation `cand` from
// `avail` (x & -x, one borrow chain over the words: v_sub_co / v_subb_co).
template <int NW>
__device__ __forceinline__ void take_lowest(uint32_t (&avail)[NW], const uint32_t (&cand)[NW]) {
  if constexpr (NW == 1) {
    avail[0] ^= cand[0] & (0u - cand[0]);
  } else if constexpr (NW == 2) {
    const uint64_t x = ((uint64_t)cand[1] << 32) | cand[0];
    const uint64_t n = 0ull - x;
    avail[0] ^= cand[0] & (uint32_t)n;
    avail[1] ^= cand[1] & (uint32_t)(n >> 32);
  } else {
    static_assert(NW <= 4, "at most 128 inner bytes in lanes");
    unsigned __int128 x = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) x |= (unsigned __int128)cand[w] << (32 * w);
    const unsigned __int128 n = (unsigned __int128)0 - x;
#pragma unroll
    for (int w = 0; w < NW; w++) avail[w] ^= cand[w] & (uint32_t)(n >> (32 * w));
  }
}

// Greedy match count (device_func.cuh:81-111) for one pair per lane, inner
// strings of at most 32*NW bytes.  The inner string is held as 8 bit-planes
// (plane k, word w: bit b set iff bit k of inner byte 32w+b is set), so the
// positions equal to an outer byte c are the AND over k of
// (plane k XNOR bit k of c), one 3-input bit operation per plane.  `win` is
// the window [i-h, i+h] of outer position i, advanced by a one-bit left
// shift per step (position 0 stays inside while i + 1 <= h: the bits shifted
// in over a block of 16 steps are the low min(16, h - base) bits of `ent`);
// `avail` holds inner positions not yet matched.  The lowest candidate is
// x & -x over the NW-word concatenation (one borrow chain), and the match
// count is S minus the positions left in `avail` at the end (each match
// takes exactly one).  No per-step liveness test is needed: cells are NUL
// padded and hold no NUL within their length, outer blocks past n_out read
// as zero, and from i >= S + h on the window lies past the inner string.
// With HIGH false no inner byte of the wave has bit 7 set, so an outer byte
// with bit 7 set matches nothing (its bit-7 mask clears `e`).
template <int NW, bool HIGH>
__device__ __forceinline__ int lane_greedy(const uint32_t (&P)[NW][8], int S, int h, int n_out,
                                           const uint8_t *outer, int n_steps) {
  uint32_t avail[NW], win[NW];
#pragma unroll
  for (int w = 0; w < NW; w++) {
    const int lo = 32 * w;
    avail[w] = S >= lo + 32 ? ~0u : (S > lo ? ((1u << (S - lo)) - 1u) : 0u);
    win[w] = h >= lo + 31 ? ~0u : (h >= lo ? (~0u >> (31 - (h - lo))) : 0u);
  }
  const uint4 *o16 = reinterpret_cast<const uint4 *>(outer);
  // the next 16 outer bytes are loaded one block ahead, so the load overlaps
  // the 16 steps of the current block
  uint4 v_next = make_uint4(0, 0, 0, 0);
  if (0 < n_out) v_next = o16[0];
  for (int base = 0; base < n_steps; base += 16) {
    const uint4 v = v_next;
    v_next = make_uint4(0, 0, 0, 0);
    if (base + 16 < n_out) v_next = o16[(base >> 4) + 1];
    const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
    const int left_in = h - base;  // steps q of this block with base + q + 1 <= h: q < left_in
    const uint32_t ent = left_in >= 16 ? 0xffffu : (left_in > 0 ? (1u << left_in) - 1u : 0u);
#pragma unroll
    for (int q = 0; q < 16; q++) {
      uint32_t sb[8];
#pragma unroll
      for (int k = 0; k < 8; k++)
        sb[k] = (uint32_t)__builtin_amdgcn_sbfe((int)vw[q >> 2], 8 * (q & 3) + k, 1);
      uint32_t cand[NW];
#pragma unroll
      for (int w = 0; w < NW; w++) {
        // a & b & ~c (0x40) and a & (b XNOR c) (0x90), bit-wise
        uint32_t e = HIGH ? (avail[w] & win[w]) : __builtin_amdgcn_bitop3_b32(avail[w], win[w], sb[7], 0x40);
#pragma unroll
        for (int k = 0; k < (HIGH ? 8 : 7); k++) e = __builtin_amdgcn_bitop3_b32(e, P[w][k], sb[k], 0x90);
        cand[w] = e;
      }
      take_lowest<NW>(avail, cand);
      // advance the window: positions shift up by one
#pragma unroll
      for (int w = NW - 1; w > 0; w--) win[w] = __builtin_amdgcn_alignbit(win[w], win[w - 1], 31);
      win[0] = (win[0] << 1) | __builtin_amdgcn_ubfe(ent, q, 1);
    }
  }
  int left = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) left += __builtin_popcount(avail[w]);
  return S - left;
}

// Words w < nwc of the row's planes; words beyond the column's stored width
// are zero (those positions lie past the inner string).
template <int NW>
__device__ __forceinline__ void load_planes(uint32_t (&P)[NW][8], const uint32_t *row_planes,
                                            int nwc) {
  const uint4 *src = reinterpret_cast<const uint4 *>(row_planes);
#pragma unroll
  for (int w = 0; w < NW; w++) {
    uint4 x = make_uint4(0, 0, 0, 0), y = make_uint4(0, 0, 0, 0);
    if (w < nwc) {
      x = src[2 * w];
      y = src[2 * w + 1];
    }
    P[w][0] = x.x; P[w][1] = x.y; P[w][2] = x.z; P[w][3] = x.w;
    P[w][4] = y.x; P[w][5] = y.y; P[w][6] = y.z; P[w][7] = y.w;
  }
}

template <int NW>
__device__ __forceinline__ int lane_matches(const uint32_t *row_planes, int nwc, int S, int T,
                                            const uint8_t *outer, bool live) {
  uint32_t P[NW][8];
  load_planes<NW>(P, row_planes, nwc);
  const int h = (S + 1) >> 1;
  const int n_out = live ? min(T, S + h) : 0;
  // uniform trip count: the largest n_out of the wave
  int n_steps = n_out;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n_steps = max(n_steps, __shfl_xor(n_steps, off));
  uint32_t hi = 0;
#pragma unroll
  for (int w = 0; w < NW; w++) hi |= P[w][7];
  if (__ballot(hi != 0))
    return lane_greedy<NW, true>(P, S, h, n_out, outer, n_steps);
  return lane_greedy<NW, false>(P, S, h, n_out, outer, n_steps);
}

// The refine test of one pair (l, r) against similarity predicate sd: the
// bound of may_pass with the 96-class counters of the refine rows, which is
// false only when JW is certainly below thr - 2e-6 (a matched pair of bytes
// is a pair of equal bytes, so m <= sum over classes of min(count_a,
// count_b) for any class partition).  A refine row holds the cell's prefix
// and length too, so a pair reads one 64-byte row per side, and both rows
// are loaded before the early outs (long or empty cells, equal first 4
// bytes).
struct RefineDesc {
  const uint32_t *hist_l, *hist_r;
  float c0, c1, c2, c3;
};

template <class SD>
__device__ __forceinline__ RefineDesc refine_desc(SD &sd) {
  return RefineDesc{sd.col_l.hist, sd.col_r.hist, sd.c0, sd.c1, sd.c2, sd.c3};
}

__device__ __forceinline__ bool hist_pass(const RefineDesc &sd, uint32_t l, uint32_t r) {
  const uint4 *ha = reinterpret_cast<const uint4 *>(sd.hist_l + (uint64_t)l * kHistWords);
  const uint4 *hb = reinterpret_cast<const uint4 *>(sd.hist_r + (uint64_t)r * kHistWords);
  const uint4 a0 = ha[0], b0 = hb[0];  // (prefix, length, 0, 0)
  uint32_t sad = 0;
  auto sad_nib = [](uint32_t a, uint32_t b, uint32_t acc) {
    acc = __builtin_amdgcn_sad_u8(a & 0x0f0f0f0fu, b & 0x0f0f0f0fu, acc);
    return __builtin_amdgcn_sad_u8((a >> 4) & 0x0f0f0f0fu, (b >> 4) & 0x0f0f0f0fu, acc);
  };
#pragma unroll
  for (int k = kHistFirstClassWord / 4; k < kHistWords / 4; k++) {
    const uint4 u = ha[k], v = hb[k];
    sad = sad_nib(u.x, v.x, sad);
    sad = sad_nib(u.y, v.y, sad);
    sad = sad_nib(u.z, v.z, sad);
    sad = sad_nib(u.w, v.w, sad);
  }
  const uint32_t la = a0.y & 0xffffu, lb = b0.y & 0xffffu;
  const uint32_t x = a0.x ^ b0.x;
  if (((a0.y | b0.y) & kLongFlag) || la == 0 || lb == 0 || x == 0) return true;
  const uint32_t mub = (la + lb - sad) >> 1;
  const uint32_t S = min(la, lb), T = max(la, lb);
  const uint32_t p = ((uint32_t)(__ffs(x) - 1)) >> 3;
  const float coef = p == 0 ? sd.c0 : (p == 1 ? sd.c1 : (p == 2 ? sd.c2 : sd.c3));
  return !((float)(mub * (S + T)) < coef * (float)(S * T));
}

// One pair per lane: the first similarity predicate of the pair's rule must
// pass hist_pass (JRs with `refine` set).  Only the first: the predicates
// are evaluated in order, and a later one failing does not spare an earlier
// one's evaluation (or its audit entry).  Rules without similarity
// predicates pass through.
//
// A work-group takes kRefineSpan consecutive entries at a time: entries of
// one filter flush are consecutive and share the rows of one tile, so the
// rows' counters are read again from the group's L2 rather than from HBM.
// Kept entries collect in a per-wave LDS buffer (one list atomic per up to
// kWaveOut entries), and filter counts in a per-wave running sum for the
// current job (one atomic when the job changes and at the end): single
// global counters taking an atomic from every wave are what limited this
// kernel otherwise.  Where a wave's 64 entries share one JR, its
// descriptors are read with scalar loads.
constexpr uint64_t kRefineSpan = 1024;

__global__ __launch_bounds__(256) void k_refine(const BatchDev b, const PairList in,
                                                const PairList out) {
  __shared__ PairEntry s_out[4][kWaveOut];
  PairEntry *wbuf = s_out[threadIdx.x >> 6];
  uint32_t wn = 0;  // entries in wbuf (wave-uniform)
  uint32_t cur_job = 0;
  unsigned long long cur_cnt = 0;  // filter survivors of cur_job not yet counted
  auto count_flush = [&]() {
    if (cur_cnt && lane_id() == 0)
      atomicAdd(b.job_cnt + (size_t)job_row(kJobFiltered, blockIdx.x & (kStatStripes - 1)) *
                                b.cnt_stride + cur_job,
                cur_cnt);
    cur_cnt = 0;
  };
  const unsigned long long n_in = *in.count;
  const unsigned long long n_use = n_in < in.cap ? n_in : in.cap;
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  for (uint64_t g0 = (uint64_t)blockIdx.x * kRefineSpan; g0 < n_use;
       g0 += (uint64_t)gridDim.x * kRefineSpan) {
    const uint64_t g_end = min(g0 + kRefineSpan, (uint64_t)n_use);
    for (uint64_t q0 = g0 + wave * 64; q0 < g_end; q0 += blockDim.x) {  // wave-uniform
      const uint64_t idx = q0 + lane;
      const bool act = idx < g_end;
      PairEntry e{0u, 0u, 0u, 0u};
      if (act) e = in.e[idx];
      const uint64_t am = __ballot(act);
      const uint32_t jr0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)e.jr);
      bool keep;
      if (__ballot(act && e.jr == jr0) == am) {  // one JR: scalar descriptor loads
        typedef const __attribute__((address_space(4))) JRDesc ConstJRDesc;
        typedef const __attribute__((address_space(4))) SimDesc ConstSimDesc;
        ConstJRDesc &d = *(ConstJRDesc *)(b.jrs + jr0);
        const uint32_t job = d.job;
        if (job != cur_job) {
          count_flush();
          cur_job = job;
        }
        cur_cnt += (unsigned long long)__popcll(am);
        const bool test = d.n_sims != 0 && d.refine;
        ConstSimDesc &sd = *(ConstSimDesc *)(b.sims + d.sim_off);
        keep = act && (!test || hist_pass(refine_desc(sd), e.l, e.r));
      } else {
        uint32_t job = 0, sim_off = 0, refine = 0;
        int32_t n_sims = 0;
        if (act) {
          const JRDesc &d = b.jrs[e.jr];
          job = d.job;
          n_sims = d.n_sims;
          sim_off = d.sim_off;
          refine = d.refine;
        }
        job_count(b, kJobFiltered, act, job);
        keep = act &&
               (n_sims == 0 || !refine || hist_pass(refine_desc(b.sims[sim_off]), e.l, e.r));
      }
      wave_append(keep, e, wbuf, wn, out, b.overflow);
    }
  }
  wave_flush(wbuf, wn, out, b.overflow);
  count_flush();
}

// A pass of the exact stage evaluates two consecutive similarity
// predicates of each pair's rule, s = a.stage and s + 1 (when the rule has
// it), 256 pairs per work-group of 8 waves: waves 0-3 run the greedy chains
// of predicate s, waves 4-7 those of predicate s + 1 for the same pairs
// (each SIMD gets one wave of each), so a pass keeps twice as many chains in
// flight as one predicate would.  The decisions then follow the reference's
// order (global_func.cuh:120-167): predicate s + 1 counts (its JW
// evaluation, its audit entry, its decision) only for pairs that passed
// predicate s.  Pairs whose shorter string of either predicate is longer
// than 128 bytes are deferred to k_exact (one wave per pair, with the
// 257-slot semantics beyond 256 bytes) for both predicates; lanes hold at
// most 4 plane words.
struct ExactSlot {
  PairEntry e;
  uint32_t job, sim_off;
  int32_t n_sims;
  uint32_t act;         // predicate s is evaluated (bit 0), and s + 1 (bit 1)
  uint32_t la[2], lb[2];
};

// The greedy match count, common prefix and lengths of one chain.
__device__ __forceinline__ void lane_chain(const SimDesc &sd, const PairEntry &e, int la, int lb,
                                           bool act, int &m, int &p) {
  const int S = min(la, lb), T = max(la, lb);
  const bool swap = la > lb;  // device_func.cuh:71-79: left stays inner on ties
  const uint8_t *ca = sd.col_l.bytes + (uint64_t)e.l * sd.col_l.width;
  const uint8_t *cb = sd.col_r.bytes + (uint64_t)e.r * sd.col_r.width;
  const uint8_t *outer = swap ? ca : cb;
  const uint32_t *planes = swap ? sd.col_r.planes + (uint64_t)e.r * sd.col_r.nwc * 8
                                : sd.col_l.planes + (uint64_t)e.l * sd.col_l.nwc * 8;
  const int nwc = (int)(swap ? sd.col_r.nwc : sd.col_l.nwc);
  p = act ? lane_prefix(ca, cb, S) : 0;
  const bool live = act && S > 0;
  const bool gt32 = __ballot(live && S > 32) != 0, gt64 = __ballot(live && S > 64) != 0;
  const bool gt96 = __ballot(live && S > 96) != 0;
  if (!gt32) m = lane_matches<1>(planes, nwc, S, T, outer, live);
  else if (!gt64) m = lane_matches<2>(planes, nwc, S, T, outer, live);
  else if (!gt96) m = lane_matches<3>(planes, nwc, S, T, outer, live);
  else m = lane_matches<4>(planes, nwc, S, T, outer, live);  // S <= 128 here
  if (!live) m = 0;
}

constexpr uint32_t kExactThreads = 512;
__global__ __launch_bounds__(kExactThreads) void k_exact_two(ExactArgs a) {
  constexpr uint32_t kPass = kExactThreads / 2;  // pairs per work-group pass
  constexpr int kClasses = 16;     // n_out classes of 16 steps
  __shared__ StageOut so;
  __shared__ ExactSlot s_slot[kPass];
  __shared__ uint32_t s_ccnt[2][kClasses];
  __shared__ uint16_t s_order[2][kPass];  // pairs of each chain half, by n_out class
  __shared__ int32_t s_m[2][kPass], s_p[2][kPass];
  const unsigned long long n_in = *a.in.count;
  const unsigned long long n_use = n_in < a.in.cap ? n_in : a.in.cap;
  const uint32_t tid = threadIdx.x;
  const int st = a.stage;
  for (uint64_t blk = (uint64_t)blockIdx.x * kPass; blk < n_use;
       blk += (uint64_t)gridDim.x * kPass) {
    if (tid == 0) {
      so.n_next = 0;
      so.n_fin = 0;
      so.fin_job = kNoJob;
    }
    if (tid < 2 * kClasses) (&s_ccnt[0][0])[tid] = 0;
    __syncthreads();
    // waves 0-1: load the pass's pairs
    uint32_t cls[2] = {0, 0}, rank[2] = {0, 0};
    if (tid < kPass) {
      const uint64_t idx = blk + tid;
      bool act = idx < n_use;
      ExactSlot sl{};
      if (act) {
        sl.e = a.in.e[idx];
        const JRDesc &d = a.b.jrs[sl.e.jr];
        sl.job = d.job;
        sl.n_sims = d.n_sims;
        sl.sim_off = d.sim_off;
      }
      if (a.count_filtered) job_count(a.b, kJobFiltered, act, sl.job);
      // a rule without similarity predicates is complete after the filter
      stage_final(a, so, act && sl.n_sims == 0, sl.e.l, sl.e.r, sl.job);
      act = act && sl.n_sims >= st;
      const bool two = act && sl.n_sims >= st + 1;
      int Smax = 0;
      for (int h = 0; h < 2; h++) {
        if (h == 0 ? act : two) {
          const SimDesc &sd = a.b.sims[sl.sim_off + st - 1 + h];
          sl.la[h] = sd.col_l.len[sl.e.l];
          sl.lb[h] = sd.col_r.len[sl.e.r];
          Smax = max(Smax, (int)min(sl.la[h], sl.lb[h]));
        }
      }
      const bool deferred = act && Smax > 128;
      list_append(deferred, sl.e, a.defer, a.b.overflow);
      act = act && !deferred;
      sl.act = act ? (two ? 3u : 1u) : 0u;
      s_slot[tid] = sl;
      for (int h = 0; h < 2; h++) {
        if (sl.act & (1u << h)) {
          const int S = (int)min(sl.la[h], sl.lb[h]), T = (int)max(sl.la[h], sl.lb[h]);
          cls[h] = (uint32_t)min(min(T, S + ((S + 1) >> 1)) >> 4, kClasses - 1);
        }
        rank[h] = atomicAdd(&s_ccnt[h][cls[h]], 1u);
      }
    }
    __syncthreads();
    if (tid < kPass) {
      for (int h = 0; h < 2; h++) {
        uint32_t off = 0;  // exclusive prefix over the classes below
        for (uint32_t c = 0; c < cls[h]; c++) off += s_ccnt[h][c];
        s_order[h][off + rank[h]] = (uint16_t)tid;
      }
    }
    __syncthreads();
    {  // the chains: half h = tid / 256 evaluates predicate st + h.  Wave w
       // of the first half takes the w-th shortest quarter of its pairs and
       // wave 4 + w the w-th longest quarter of the second half's (both run
       // on SIMD w), so that every SIMD gets a similar share of steps.
      const int h = (int)(tid / kPass);
      const uint32_t k = tid % kPass;
      const uint32_t i = s_order[h][h ? kPass - 1 - k : k];
      const ExactSlot &sl = s_slot[i];
      const bool act = (sl.act >> h) & 1u;
      SimDesc sd{};
      if (act) sd = a.b.sims[sl.sim_off + st - 1 + h];
      int m = 0, p = 0;
      lane_chain(sd, sl.e, (int)sl.la[h], (int)sl.lb[h], act, m, p);
      s_m[h][i] = m;
      s_p[h][i] = p;
    }
    __syncthreads();
    if (tid < kPass) {  // decisions in the reference's order
      const ExactSlot sl = s_slot[tid];
      const bool act = sl.act & 1u;
      SimDesc sd{};
      if (act) sd = a.b.sims[sl.sim_off + st - 1];
      const int S0 = (int)min(sl.la[0], sl.lb[0]), T0 = (int)max(sl.la[0], sl.lb[0]);
      const bool pass0 = decide(a, sl.e, sl.job, sd.thr, sd.slot, s_m[0][tid], S0, T0,
                                s_p[0][tid], act);
      job_count(a.b, kJobJw, act, sl.job);
      const bool eval1 = pass0 && (sl.act & 2u);  // predicate st + 1, reached
      SimDesc sd1{};
      if (eval1) sd1 = a.b.sims[sl.sim_off + st];
      const int S1 = (int)min(sl.la[1], sl.lb[1]), T1 = (int)max(sl.la[1], sl.lb[1]);
      const bool pass1 = decide(a, sl.e, sl.job, sd1.thr, sd1.slot, s_m[1][tid], S1, T1,
                                s_p[1][tid], eval1);
      job_count(a.b, kJobJw, eval1, sl.job);
      // the rule's last predicate decides; earlier ones send the pair on
      const int last_done = (sl.act & 2u) ? st + 1 : st;
      const bool passed = (sl.act & 2u) ? pass1 : pass0;
      const bool done = act && sl.n_sims == last_done;
      stage_final(a, so, passed && done, sl.e.l, sl.e.r, sl.job);
      stage_put(passed && !done && act, sl.e, so.next, so.n_next);
    }
    stage_flush(a, so);
  }
}

// One wave per pair (scalar greedy); used for pairs deferred by
// k_exact_two, i.e. shorter strings of more than 256 bytes.
__global__ __launch_bounds__(256) void k_exact(ExactArgs a) {
  const uint32_t lane = lane_id();
  const unsigned long long n_in = *a.in.count;
  const unsigned long long n_use = n_in < a.in.cap ? n_in : a.in.cap;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t n_waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t idx = wave; idx < n_use; idx += n_waves) {
    const PairEntry e = a.in.e[idx];
    const JRDesc &d = a.b.jrs[e.jr];
    // predicates stage and stage + 1 (as in k_exact_two), in order
    bool pass = true;
    int stage = a.stage;
    const int last = min(d.n_sims, a.stage + 1);
    for (; stage <= last && pass; stage++) {
      const SimDesc sd = a.b.sims[d.sim_off + stage - 1];
      const uint8_t *ca = sd.col_l.bytes + (uint64_t)e.l * sd.col_l.width;
      const uint8_t *cb = sd.col_r.bytes + (uint64_t)e.r * sd.col_r.width;
      const int la = sd.col_l.len[e.l], lb = sd.col_r.len[e.r];
      const int S = min(la, lb), T = max(la, lb);
      // device_func.cuh:71-79: the left cell stays inner unless strictly longer.
      const bool swap = la > lb;
      const uint8_t *inner = swap ? cb : ca;
      const uint8_t *outer = swap ? ca : cb;
      const uint32_t outer_width = swap ? sd.col_l.width : sd.col_r.width;
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
      pass = decide(a, e, d.job, sd.thr, sd.slot, m, S, T, p, lane == 0);
      pass = __shfl(pass ? 1 : 0, 0) != 0;  // lane 0's decision, for the wave
      job_count(a.b, kJobJw, lane == 0, d.job);
    }
    const bool done = last == d.n_sims;
    finalize(a, pass && done && lane == 0, e.l, e.r, d.job);
    list_append(pass && !done && lane == 0, e, a.next, a.b.overflow);
  }
}

// Copies the final pairs and the audit entries into per-job ranges.
__global__ void k_scatter(const BatchDev b, const PairList fin, const uint64_t *__restrict__ out_off,
                          unsigned long long *fill, uint32_t *__restrict__ out_pairs,
                          unsigned long long out_cap, const AuditList au,
                          const uint64_t *__restrict__ audit_off,
                          unsigned long long *audit_fill, hb_audit_entry *__restrict__ out_audit,
                          unsigned long long audit_cap) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t base0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
  const unsigned long long n = min(*fin.count, fin.cap);
  // A batch of one job (cnt_stride == 2) keeps the list order: no slots to
  // hand out, so no atomics.
  const bool one_job = b.cnt_stride == 2;
  for (uint64_t q0 = base0; q0 < n; q0 += stride) {
    const uint64_t idx = q0 + lane_id();
    const bool act = idx < n;
    PairEntry e{0u, 0u, 0u, 0u};
    if (act) e = fin.e[idx];
    const unsigned long long k = one_job ? idx : job_slot(act, e.jr, fill);
    if (act) {
      const unsigned long long pos = out_off[e.jr] + k;
      if (pos < out_cap) {
        reinterpret_cast<uint2 *>(out_pairs)[pos] = make_uint2(e.l, e.r);  // one 8-byte store
      } else {
        *b.overflow = 1ull;
      }
    }
  }
  const unsigned long long na = min(*au.count, au.cap);
  for (uint64_t q0 = base0; q0 < na; q0 += stride) {
    const uint64_t idx = q0 + lane_id();
    const bool act = idx < na;
    uint32_t job = 0;
    if (act) job = au.e[idx].job;
    const unsigned long long k = one_job ? idx : job_slot(act, job, audit_fill);
    if (act) {
      const unsigned long long pos = audit_off[job] + k;
      if (pos < audit_cap)
        out_audit[pos] = au.e[idx].e;
      else
        *b.overflow = 1ull;
    }
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

void launch_build_hist(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint32_t *hist,
                       hipStream_t stream) {
  if (n_rows <= 0) return;
  hipLaunchKernelGGL(k_build_hist, dim3(grid_for((uint64_t)n_rows, 256)), dim3(256), 0, stream,
                     bytes, width, n_rows, hist);
}

void launch_refine(BatchDev b, PairList in, PairList out, int n_cu, hipStream_t stream) {
  const unsigned cus = (unsigned)(n_cu > 0 ? n_cu : 256);
  hipLaunchKernelGGL(k_refine, dim3(cus * 8), dim3(256), 0, stream, b, in, out);
}

void launch_build_planes(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint32_t nwc,
                         uint32_t *planes, hipStream_t stream) {
  if (n_rows <= 0) return;
  hipLaunchKernelGGL(k_build_planes, dim3(grid_for((uint64_t)n_rows, 256)), dim3(256), 0, stream,
                     bytes, width, n_rows, nwc, planes);
}

void launch_build_col(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint32_t nwc,
                      uint16_t *len, SigRow *sig, uint32_t *hist, uint32_t *planes,
                      hipStream_t stream) {
  if (n_rows <= 0) return;
  const uint32_t groups = (uint32_t)((n_rows + 63) / 64);
  if (hist)
    k_build_col<true><<<groups, 256, 0, stream>>>(bytes, width, n_rows, nwc, len, sig, hist, planes);
  else
    k_build_col<false><<<groups, 256, 0, stream>>>(bytes, width, n_rows, nwc, len, sig, hist, planes);
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

// ---- scratch of the no-allocation entry point --------------------------
//
// hb_block_noalloc carves all of a job's working memory (counters,
// descriptors, join groups, pair lists, dedupe sets, scan temporaries) from
// this array instead of allocating: each GPU the library runs on holds its
// own copy for the life of the process.  Host code reaches it only through
// the device address noalloc_scratch returns (the array's name is a
// host-side stand-in); calls that overlap on one GPU share it and race.
}  // namespace hb
static __device__ __align__(256) unsigned char scratch[hb::kNoAllocScratchBytes];
namespace hb {

hipError_t noalloc_scratch(void **ptr) {
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

void launch_dict_code(const uint8_t *cells, uint32_t width, uint32_t fixed_len, uint32_t n,
                      DictDev d, uint32_t *code, uint32_t *fb_rows,
                      unsigned long long *fb_count, unsigned long long *status,
                      hipStream_t stream) {
  const uint32_t blocks = std::max<uint32_t>(1, (n + 255) / 256);
  k_dict_insert<<<blocks, 256, 0, stream>>>(cells, width, fixed_len, n, d, code, fb_count);
  // Work-groups of 1024: each group's fence before its done count writes
  // back its L2, so fewer groups end the launch sooner.
  const uint32_t vblocks = std::max<uint32_t>(1, (n + 1023) / 1024);
  k_dict_verify<<<vblocks, 1024, 0, stream>>>(cells, width, fixed_len, n, d, code, fb_rows,
                                              fb_count, status);
}

void launch_compose_keys(KeyColumns cols, uint32_t n, uint8_t *cells, uint32_t width,
                         hipStream_t stream) {
  if (!n) return;
  k_compose_keys<<<(n + 255) / 256, 256, 0, stream>>>(cols, n, cells, width);
}

void launch_mask_keys(KeyColumns cols, uint32_t n, uint32_t *code, hipStream_t stream) {
  if (!n) return;
  k_mask_keys<<<(n + 255) / 256, 256, 0, stream>>>(cols, n, code);
}

void launch_unpack_rows(const uint8_t *rows, uint64_t row_bytes, uint32_t n_rows,
                        const UnpackColumns &cols, const UnpackColumn *cols_dev, uint32_t n_cols,
                        uint32_t rows_per_group, int n_cu, hipStream_t stream) {
  if (!n_rows || !n_cols) return;
  const uint32_t span = (uint32_t)(rows_per_group * row_bytes);
  const size_t lds = unpack_lds_bytes(span, (size_t)rows_per_group * n_cols);
  const uint32_t n_blocks = (n_rows + rows_per_group - 1) / rows_per_group;
  const uint32_t grid = std::min<uint32_t>(n_blocks, (uint32_t)n_cu * 8);
  if ((((uintptr_t)rows) & 15u) == 0 && (row_bytes & 15u) == 0)
    k_unpack_rows<true><<<grid, 256, lds, stream>>>(rows, row_bytes, n_rows, cols, cols_dev,
                                                     n_cols, rows_per_group);
  else
    k_unpack_rows<false><<<grid, 256, lds, stream>>>(rows, row_bytes, n_rows, cols, cols_dev,
                                                      n_cols, rows_per_group);
}

void launch_layout_small(const LayoutSmall &a, hipStream_t stream) {
  if (!a.n) return;
  const uint32_t blocks = (a.n + 255) / 256;
  const size_t lds = (a.sig ? 2 : 1) * a.n_codes * 256 * sizeof(uint32_t);
  const uint32_t groups = (a.n + kLayoutRowsPerGroup - 1) / kLayoutRowsPerGroup;
  if (a.n_codes <= kLayoutLdsCodes) {
    k_lsmall_count_lds<<<groups, 1024, lds, stream>>>(a);
    k_lsmall_place<<<groups, 1024, lds, stream>>>(a);
  } else {
    k_lsmall_count<<<groups, 1024, 0, stream>>>(a);
    k_lsmall_scatter<<<blocks, 256, 0, stream>>>(a);
  }
}

void launch_layout_keys(const uint32_t *code, const uint16_t *len, uint32_t n,
                        uint64_t code_bound, uint32_t *bad, uint64_t *keys, hipStream_t stream) {
  if (!n) return;
  k_layout_keys<<<(n + 255) / 256, 256, 0, stream>>>(code, len, n, code_bound, bad, keys);
}

void launch_allow_table(float c0, float c3, uint32_t *tab, hipStream_t stream) {
  hipLaunchKernelGGL(k_allow_table, dim3(grid_for(kAllowRows * kAllowCols, 256)), dim3(256), 0,
                     stream, c0, c3, tab);
}

void launch_layout_keys_fb(const uint32_t *code, const SigRow *sig, const uint16_t *len,
                           uint32_t n, uint64_t code_bound, uint32_t *bad, uint64_t *keys,
                           hipStream_t stream) {
  if (!n) return;
  k_layout_keys_fb<<<(n + 255) / 256, 256, 0, stream>>>(code, sig, len, n, code_bound, bad, keys);
}

void launch_gather_sig(const uint32_t *perm, uint32_t n, const SigRow *sig, SigRow *jsig,
                       hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(k_gather_sig, dim3(grid_for(n, 256)), dim3(256), 0, stream, perm, n, sig,
                     jsig);
}

void launch_run_flags(const uint64_t *sorted_keys, uint32_t n, uint32_t *flags,
                      hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(k_run_flags, dim3(grid_for(n, 256)), dim3(256), 0, stream, sorted_keys, n,
                     flags);
}

void launch_layout_runs(const uint64_t *sorted_keys, const uint32_t *flags,
                        const uint32_t *run_id, uint32_t n, uint32_t *run_code,
                        uint32_t *run_off, uint32_t *n_runs, uint32_t *n_runs_host,
                        hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(k_layout_runs, dim3(grid_for(n, 256)), dim3(256), 0, stream, sorted_keys,
                     flags, run_id, n, run_code, run_off, n_runs, n_runs_host);
}

void launch_batch_prologue(BatchDev b, GroupArrays g, const void *hdesc_dev, void *ddesc,
                           size_t desc_bytes, unsigned long long *cnt, size_t cnt_words,
                           uint64_t set_words, hipStream_t stream) {
  hipLaunchKernelGGL(k_batch_prologue, dim3(1), dim3(1024), 0, stream, b, g,
                     (const uint4 *)hdesc_dev, (uint4 *)ddesc, (uint32_t)((desc_bytes + 15) / 16),
                     cnt, (uint32_t)cnt_words, set_words);
}

void launch_batch_join(BatchDev b, GroupArrays g, hipStream_t stream) {
  hipLaunchKernelGGL(k_batch_join, dim3(grid_for((uint64_t)g.n + 1, 256)), dim3(256), 0, stream,
                     b, g);
}

void launch_filter(BatchDev b, GroupArrays g, uint32_t g_begin, uint32_t g_end, int variant,
                   PairList out, FilterShape shape, int n_cu, hipStream_t stream) {
  if (g_end <= g_begin) return;
  // The tile grid launches grid_per_cu work-groups per CU (16 by default,
  // two rounds of the 8 that fit at once), each taking work items in a
  // grid-stride loop; surplus groups find no item and exit.
  const unsigned cus = (unsigned)(n_cu > 0 ? n_cu : 256);
  // The last cus * 8 work-groups of the launch take the small groups.
  const unsigned tile_grid = cus * shape.grid_per_cu, grid = tile_grid + cus * 8;
#define HB_FILTER(SIM, TAB)                                                                  \
  do {                                                                                       \
    if (shape.rows_per_lane == 4)                                                            \
      hipLaunchKernelGGL((k_filter_tiles<SIM, TAB, 4, 1>), dim3(grid), dim3(256), 0, stream, \
                         b, g, g_begin, g_end, out, tile_grid);                              \
    else if (shape.rows_per_lane == 2)                                                       \
      hipLaunchKernelGGL((k_filter_tiles<SIM, TAB, 2, 2>), dim3(grid), dim3(256), 0, stream, \
                         b, g, g_begin, g_end, out, tile_grid);                              \
    else                                                                                     \
      hipLaunchKernelGGL((k_filter_tiles<SIM, TAB, 1, 2>), dim3(grid), dim3(256), 0, stream, \
                         b, g, g_begin, g_end, out, tile_grid);                              \
  } while (0)
  if (variant == kFilterSimTable)
    HB_FILTER(true, true);
  else if (variant == kFilterSim)
    HB_FILTER(true, false);
  else
    HB_FILTER(false, false);
#undef HB_FILTER
}

void launch_exact(BatchDev b, PairList in, int stage, PairList next, PairList final_,
                  PairList defer, AuditList audit, bool count_filtered, bool may_defer,
                  uint2 *direct_pairs, hb_audit_entry *direct_audit, int n_cu,
                  hipStream_t stream) {
  const unsigned cus = (unsigned)(n_cu > 0 ? n_cu : 256);
  ExactArgs a{b,      in,    stage,        count_filtered ? 1 : 0, next, final_, defer, audit,
              direct_pairs, direct_audit};
  hipLaunchKernelGGL(k_exact_two, dim3(cus * 4), dim3(kExactThreads), 0, stream, a);
  if (!may_defer) return;
  ExactArgs d{b, defer, stage, 0, next, final_, defer, audit, direct_pairs, direct_audit};
  d.defer.cap = 0;  // deferred pairs are not deferred again
  hipLaunchKernelGGL(k_exact, dim3(cus * 4), dim3(256), 0, stream, d);
}

void launch_scatter(BatchDev b, PairList final_, const uint64_t *out_off,
                    unsigned long long *fill, uint32_t *out_pairs, unsigned long long out_cap,
                    AuditList audit, const uint64_t *audit_off, unsigned long long *audit_fill,
                    hb_audit_entry *out_audit, unsigned long long audit_cap, int n_cu,
                    hipStream_t stream) {
  const unsigned cus = (unsigned)(n_cu > 0 ? n_cu : 256);
  hipLaunchKernelGGL(k_scatter, dim3(cus * 2), dim3(256), 0, stream, b, final_, out_off, fill,
                     out_pairs, out_cap, audit, audit_off, audit_fill, out_audit, audit_cap);
}

void launch_jw_from_counts(const int32_t *mstp, int64_t n, double *out, double *out_fused,
                           hipStream_t stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_jw_from_counts, dim3(grid_for((uint64_t)n, 256)), dim3(256), 0, stream,
                     mstp, n, out, out_fused);
}

}  // namespace hb
