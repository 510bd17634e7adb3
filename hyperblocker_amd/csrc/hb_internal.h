// Internal layout shared by the host runtime and the gfx950 kernels.
//
// Device-resident table, per column c (structure of arrays):
//   bytes[c] : n_rows x width[c] bytes, each cell NUL padded (every byte after
//              the first NUL is zero), width a multiple of 16 and at least
//              the longest cell + 1
//   len[c]   : uint16 byte length of each cell (up to its first NUL)
//   sig[c]   : 48-byte filter signature per row (SigRow), built on first use
//              by a similarity predicate
//   code[c]  : uint32 dictionary code per row, built on first use by an
//              equality predicate; equal codes <=> equal byte strings
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#else
#include <hip/hip_runtime_api.h>
#endif

namespace hb {

// Filter signature of one cell.
//   h[8] : 32 byte-sized counters, counter b holds #bytes c with (c & 31) == b
//          (counter b lives in byte (b & 3) of word b >> 2)
//   pre  : first 4 bytes of the cell, NUL padded, little endian
//   len  : bits 0-15 length; bit 31 set when the length exceeds 255 (the
//          counters could saturate, so the filter passes such pairs on)
//   rlen : 1/length rounded to float when 1 <= length <= 255, else +inf; an
//          infinite term makes the bound's left-hand side infinite (or NaN),
//          so the pair always passes
struct __attribute__((aligned(16))) SigRow {
  uint32_t h[8];
  uint32_t pre;
  uint32_t len;
  float rlen;
  uint32_t pad;
};
static_assert(sizeof(SigRow) == 48, "SigRow layout");

constexpr uint32_t kLongFlag = 0x80000000u;
constexpr int kMaxEq = 4;        // equality predicates per rule beyond the join key

// Bit-planes of a similarity column (positions 0..32*nwc-1, nwc <= 8):
// uint32 planes[row][w][k], bit b of word (w, k) is bit k of byte 32w+b.
constexpr int kMaxPlaneWords = 8;

struct DevColumn {
  const uint8_t *bytes;
  const uint16_t *len;
  uint32_t width;
  const uint32_t *planes;
  uint32_t nwc;
};

// Join groups.  Run g of equal key in the sorted left keys: left positions
// [l_off, l_off + l_cnt) of perm_l, right positions [r_off, r_off + r_cnt)
// of perm_r.  Arrays hold n_l + 1 entries; entries at and beyond *n_runs are
// zero so the exclusive scans give totals at index n_l.
struct GroupArrays {
  uint32_t *n_runs;        // [1]
  uint32_t *l_off, *l_cnt, *r_off, *r_cnt;
  uint64_t *small_pairs;   // pairs of groups with <= small_pairs_max pairs, else 0
  uint64_t *pair_off;      // exclusive scan of small_pairs
  uint64_t *items;         // work items of the other groups, else 0
  uint64_t *item_off;      // exclusive scan of items
  unsigned long long *joined;  // += pairs sharing the join key
  uint32_t cap;            // n_l: array entries - 1
  uint32_t tile_a;         // rows of the larger side per work item (256 * rows per lane)
  uint32_t tile_b;         // rows of the smaller side per work item
  uint32_t small_pairs_max;  // groups with at most this many pairs take the list path
};

// Launch shape of the filter stage (chosen per device by the autotuner).
struct FilterShape {
  int rows_per_lane;   // 1, 2 or 4
  uint32_t tile_b;     // multiple of 256
  unsigned grid_per_cu;
};

struct EqCheck {  // equality predicates checked per pair (beyond the join key)
  int32_t n;
  const uint32_t *code_l[kMaxEq];
  const uint32_t *code_r[kMaxEq];
};

struct StageOut {  // pair list output with capacity
  uint32_t *pairs;                  // 2 * cap uint32
  unsigned long long *count;
  unsigned long long cap;
  unsigned long long *overflow;     // set to 1 when count exceeded cap
};

struct AuditOut {
  void *entries;  // hb_audit_entry[cap]
  unsigned long long *count;
  unsigned long long cap;
};

// Kernel launch wrappers (hb_kernels.hip).  All are asynchronous on stream.
void launch_build_sig(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint16_t *len,
                      SigRow *sig, hipStream_t stream);
void launch_build_planes(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint32_t nwc,
                         uint32_t *planes, hipStream_t stream);
void launch_iota(uint32_t *out, uint32_t n, hipStream_t stream);
// Join order of a table for one rule: keys (code << 8) | min(length, 255)
// (code or length may be absent: null pointer), sorted with the row ids.
void launch_layout_keys(const uint32_t *code, const uint16_t *len, uint32_t n, uint64_t *keys,
                        hipStream_t stream);
void launch_gather_sig(const uint32_t *perm, uint32_t n, const SigRow *sig, SigRow *jsig,
                       hipStream_t stream);
// Runs of equal code (key >> 8) in the sorted left keys, matched with the
// sorted right keys.
void launch_run_flags(const uint64_t *sorted_keys, uint32_t n, uint32_t *flags,
                      hipStream_t stream);
void launch_run_info(const uint64_t *sorted_l, const uint32_t *flags, const uint32_t *run_id,
                     uint32_t n_l, const uint64_t *sorted_r, uint32_t n_r, GroupArrays g,
                     hipStream_t stream);
void launch_single_group(uint32_t n_l, uint32_t n_r, GroupArrays g, hipStream_t stream);

// Signatures and extra equality codes copied into join order (element i of
// the outputs belongs to row perm[i]; codes kMaxEq per row).  side 0 = left.
void launch_gather_join(const uint32_t *perm, uint32_t n, const SigRow *sig, SigRow *jsig,
                        EqCheck eq, int side, uint32_t *jcode, hipStream_t stream);

// First stage over join groups: remaining equalities + bound filter of the
// first similarity predicate (has_sim = 0: every equal pair is emitted).
// Inputs are in join order (launch_gather_join).
void launch_filter_groups(GroupArrays g, const uint32_t *perm_l, const uint32_t *perm_r,
                          const SigRow *jsig_l, const SigRow *jsig_r, const uint32_t *jcode_l,
                          const uint32_t *jcode_r, int n_eq, int has_sim, const float coef[4],
                          StageOut out, FilterShape shape, int n_cu, hipStream_t stream);

// Exact Jaro-Winkler on a pair list (count read from the device): one pair
// per lane, then one wave per pair for the pairs moved to `defer` (shorter
// string longer than 256 bytes).
void launch_exact(const uint32_t *pairs, const unsigned long long *n_pairs,
                  unsigned long long in_cap, DevColumn col_l, DevColumn col_r, float threshold,
                  int32_t slot, StageOut out, AuditOut audit, unsigned long long *jw_count,
                  StageOut defer, int n_cu, hipStream_t stream);

// Append a pair list to the job output, dropping pairs already present
// (device hash set keyed by (l, r); empty slots hold ~0).
void launch_merge_unique(const uint32_t *pairs, const unsigned long long *n_pairs,
                         unsigned long long in_cap, unsigned long long *set,
                         unsigned long long set_cap, StageOut out, int n_cu, hipStream_t stream);
void launch_fill_u64(unsigned long long *p, unsigned long long n, unsigned long long v,
                     hipStream_t stream);

// Jaro-Winkler from counts, one thread per (m, S, T, p) tuple (tests).
void launch_jw_from_counts(const int32_t *mstp, int64_t n, double *out, double *out_fused,
                           hipStream_t stream);

// hb_sort.hip: rocPRIM wrappers.  temp may be null to query temp_bytes.
hipError_t sort_pairs_u32(void *temp, size_t &temp_bytes, const uint32_t *keys_in,
                          uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                          uint32_t n, hipStream_t stream);
// Stable radix sort of (key, value) pairs on key bits [0, end_bit).
hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const uint64_t *keys_in,
                          uint64_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                          uint32_t n, int end_bit, hipStream_t stream);
hipError_t exscan_u32(void *temp, size_t &temp_bytes, const uint32_t *in, uint32_t *out,
                      uint32_t n, hipStream_t stream);
hipError_t exscan_u64(void *temp, size_t &temp_bytes, const uint64_t *in, uint64_t *out,
                      uint32_t n, hipStream_t stream);

}  // namespace hb
