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
// and, per list of two or more equality columns used together by a rule, a
// composite code per row (equal <=> every listed cell equal).
//
// Work is submitted in batches of jobs (a plain job is a batch of one).  A
// job-rule (JR) is one rule of one job's plan; every kernel of a batch
// covers all of its JRs, so a batch costs the same number of launches
// whatever its number of jobs.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#else
#include <hip/hip_runtime_api.h>
#endif

#include "../../include/hyperblocker_amd.h"

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
static_assert(__builtin_offsetof(SigRow, pre) == 32, "SigRow: pre opens the third 16-byte word");

constexpr uint32_t kLongFlag = 0x80000000u;

// The refine pass's row of a similarity column, 64 bytes: word 0 the
// cell's first 4 bytes and word 1 its length (with kLongFlag), as in its
// SigRow; words 2-3 zero; words 4-15 byte-class counters.  The class of a
// byte is byte & 127, except that classes 0-31 are counted in class + 32
// (any partition of the bytes into classes gives a sound bound, and this
// one frees the first 16 bytes).  Counters are saturated at 15 and kept as
// nibbles: word k holds classes 8k .. 8k + 7, class 8k + e (e < 4) in the
// low nibble of byte e and class 8k + 4 + e in its high nibble, so that
// (w & 0x0f0f0f0f) and ((w >> 4) & 0x0f0f0f0f) are u8 counters for
// v_sad_u8.  Saturation only lowers |a - b| per class, so the bound built
// on these counters stays an upper bound.  Counters of cells longer than
// 255 bytes may be wrong; such cells carry kLongFlag and no bound rejects
// them.
constexpr int kHistWords = 16;
constexpr int kHistFirstClassWord = 4;
// The u8 counters the builders accumulate per row (class k in byte k & 3 of
// word k >> 2).
constexpr int kClassWords = 32;

// Bit-planes of a similarity column (positions 0..32*nwc-1, nwc <= 8):
// uint32 planes[row][w][k], bit b of word (w, k) is bit k of byte 32w+b.
constexpr int kMaxPlaneWords = 8;

struct DevColumn {
  const uint8_t *bytes;
  const uint16_t *len;
  uint32_t width;
  const uint32_t *planes;
  uint32_t nwc;
  const SigRow *sig;      // per row
  const uint32_t *hist;   // kHistWords per row
};

// One side (0 = left, 1 = right) of a JR, in its join order: position i is
// row perm[i].  Runs of equal join key (the code of the rule's equality
// column, or the composite code of its equality columns): run g covers
// positions [run_off[g], run_off[g + 1]) and has code run_code[g] (codes
// ascending; a rule without equalities has a single run of code 0).
struct SideDesc {
  const uint32_t *perm;
  const SigRow *jsig;      // signatures of the rule's first similarity column, or null
  // The same rows ordered by (code, first byte, length): within a join group
  // the rows are those of perm/jsig, ordered by their first byte.  Equal to
  // perm/jsig when the rule has no similarity predicate.
  const uint32_t *perm_fb;
  const SigRow *jsig_fb;
  const uint32_t *run_code;
  const uint32_t *run_off;  // n_runs + 1 entries
  // Runs on the host side: the count, or (a layout still being built) an
  // upper bound of it that sizes the JR's groups; the count itself is at
  // *n_runs_dev (the join leaves groups past it empty).
  uint32_t n_runs;
  uint32_t pad;
  const uint32_t *n_runs_dev;
  // Set by the layout's build when a row's code was not below the code
  // count the build assumed (a value coded since, or one the host codes):
  // the layout is then rebuilt and the batch run again.
  const uint32_t *bad;
};

struct JRDesc {
  SideDesc s[2];
  float c0, c1, c2, c3;  // bound-filter coefficients of the first similarity
  uint32_t job;
  int32_t n_sims;
  uint32_t sim_off;      // first SimDesc of this JR
  uint32_t multi;        // the job has several rules: its final pairs are deduplicated
  uint32_t refine;       // the refine pass tests the first similarity predicate
  uint32_t pad;
  // Allowance table of (2*c0, 2*c3), or null: entry (lb, la) at
  // allow_tab[2 * (lb * kAllowCols + la)] holds the negated allowances
  // (class 0, class 3) for an A row of length la and a B row of length lb
  // (lengths 1..255; column 0 for an empty or over-long cell, column 256
  // for a padding A row).
  const uint32_t *allow_tab;
};
constexpr uint32_t kAllowCols = 257, kAllowRows = 256;

struct SimDesc {
  DevColumn col_l, col_r;
  float thr;
  int32_t slot;
  float c0, c1, c2, c3;  // bound coefficients of thr (as in JRDesc)
};

// Per-job counters, structure of arrays: counter k of job j lives at
// job_cnt[k * cnt_stride + j], cnt_stride = jobs + 1 (the extra entry stays
// zero, so an exclusive scan of a row gives its total at index jobs).
enum : int {
  kJobJoined = 0,    // candidate pairs (sharing the join key), over rules
  kJobFiltered = 1,  // pairs the bound filter kept
  kJobJw = 2,        // exact JW evaluations
  kJobMatches = 3,   // distinct output pairs
  kJobAudit = 4,     // audit entries
  kJobCounters = 5,
};

// The statistics-only counters (joined, filtered, JW) have kStatStripes
// rows each, chosen by work-group, so that thousands of waves adding to one
// job's count do not all wait on one address; the host sums the rows.  Rows
// beyond the first kJobCounters hold stripes 1.. of those three counters.
constexpr int kStatStripes = 16;
constexpr int kJobRows = kJobCounters + 3 * (kStatStripes - 1);
__host__ __device__ inline int job_row(int k, int stripe) {
  return (k <= kJobJw && stripe != 0) ? kJobCounters + k * (kStatStripes - 1) + stripe - 1 : k;
}

// Per-job hash sets deduplicating the output of jobs with several rules:
// job j owns slots [off[j], off[j] + cap[j]) (cap a power of two, 0 for jobs
// with one rule); empty slots hold ~0.
struct JobSets {
  unsigned long long *slots;
  const unsigned long long *off;
  const unsigned long long *cap;
};

struct BatchDev {
  const JRDesc *jrs;
  const uint32_t *jr_group_off;  // n_jr + 1: first group of each JR
  uint32_t n_jr;
  uint32_t cnt_stride;
  const SimDesc *sims;
  unsigned long long *job_cnt;   // kJobRows * cnt_stride
  JobSets sets;
  unsigned long long *overflow;  // set to 1 when any list or set is full
  unsigned long long *set_overflow;  // set to 1 when a job's dedupe set is full
  unsigned long long *invalid;   // a JR's layout is flagged bad (SideDesc::bad)
};

// Join groups of a batch: group gi pairs left positions [l_off, l_off + l_cnt)
// with right positions [r_off, r_off + r_cnt) of its JR.  The arrays hold
// n + 1 entries; entry n is zero so that exclusive scans give totals there.
struct GroupArrays {
  uint32_t n;
  uint32_t *l_off, *l_cnt, *r_off, *r_cnt;
  uint32_t *jr;            // the group's JR
  uint64_t *small_pairs;   // pairs of groups with <= small_pairs_max pairs, else 0
  uint64_t *items;         // work items of the other groups, else 0
  uint64_t *item_off;      // exclusive scan of items
  uint32_t tile_a;         // rows of the larger side per work item (256 * rows per lane)
  uint32_t tile_b;         // rows of the smaller side per work item
  uint32_t small_pairs_max;  // groups with at most this many pairs run one wave per group
};

// Launch shape of the filter stage (chosen per device by the autotuner).
struct FilterShape {
  int rows_per_lane;   // 1, 2 or 4
  uint32_t tile_b;     // multiple of 64
  unsigned grid_per_cu;
  uint32_t small_pairs_max = 2048;  // groups up to this many pairs run one wave per group
};

// Pair list entry: rows l, r of JR jr (in the final list: of job jr).
struct PairEntry {
  uint32_t l, r, jr, pad;
};

struct PairList {
  PairEntry *e;
  unsigned long long *count;
  unsigned long long cap;
};

struct AuditRec {
  hb_audit_entry e;
  uint32_t job, pad;
};

struct AuditList {
  AuditRec *e;
  unsigned long long *count;
  unsigned long long cap;
};

// Dictionary of a context in device memory: equality columns (and the code
// tuples of composite keys) are coded here, so that equal codes mean equal
// byte strings across every table of the context.  Open addressing keyed by
// a 64-bit hash of the cell; the first row to claim a slot gives the value
// the next code (codes are dense, so join keys stay narrow) and copies its
// bytes to the arena as the slot's
// representative, and every row is then compared byte for byte with the
// representative of its slot: a row that differs (a hash collision), or that
// found no slot within max_probe probes, or whose slot has no representative
// (arena full), gets no code here and is coded by the host's exact dictionary
// instead (codes from kFallbackCodeBase up).
struct DictDev {
  unsigned long long *keys;        // slots: hash, or kDictEmpty
  unsigned long long *reps;        // slots: arena offset << 16 | chunks, or kDictNoRep
  uint32_t *slot_code;             // slots: the value's code (dense: in order of first sight)
  uint32_t *n_codes;               // codes handed out
  uint8_t *arena;
  unsigned long long *arena_used;  // bytes handed out (may exceed arena_cap)
  unsigned long long arena_cap;
  uint32_t mask;                   // slots - 1 (slots a power of two, <= 2^30)
  uint32_t max_probe;
  uint32_t hash_bits;              // 64; fewer only to force collisions in tests
  uint32_t pad;
  uint32_t *done;                  // work-groups of the running k_dict_verify done (0 at rest)
};
constexpr unsigned long long kDictEmpty = ~0ull;
constexpr unsigned long long kDictNoRep = ~0ull;
constexpr uint32_t kNoCode = 0xffffffffu;
constexpr uint32_t kFallbackCodeBase = 0x80000000u;
constexpr int kMaxKeyColumns = 16;  // equality columns of one composite key

struct KeyColumns {
  const uint32_t *code[kMaxKeyColumns];
  uint32_t k;
};

// One column of a row-major table (the reference's SerializedTable) to
// unpack into a column of NUL-padded cells.
struct UnpackColumn {
  uint8_t *out;      // n_rows x width bytes
  uint64_t offset;   // byte offset of the cell in a row
  uint32_t size;     // bytes of the cell in a row (col_size)
  uint32_t width;    // multiple of 16, > size
};

// Working memory of hb_block_noalloc: a file-scope device array of this
// many bytes per GPU (hb_kernels.hip); noalloc_scratch gives its device
// address on the current device (hipGetSymbolAddress).
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
hipError_t noalloc_scratch(void **ptr);

// Kernel launch wrappers (hb_kernels.hip).  All are asynchronous on stream.

// Dictionary codes of n cells (each `width` bytes, 16-byte aligned, NUL
// padded: a cell ends at its first NUL, or has fixed_len bytes when
// fixed_len > 0).  Rows left to the host get kNoCode in `code` and are listed
// in fb_rows[0 .. *fb_count) (fb_rows holds n entries; the kernels zero
// *fb_count first).  When done, the last work-group writes (that count, the
// codes handed out) to status[0..1] (pinned host memory, device address).
void launch_dict_code(const uint8_t *cells, uint32_t width, uint32_t fixed_len, uint32_t n,
                      DictDev d, uint32_t *code, uint32_t *fb_rows,
                      unsigned long long *fb_count, unsigned long long *status,
                      hipStream_t stream);
// Key cells of a composite key: byte 0xC0 + k, then the k column codes
// (little endian), zero padded to width (>= 1 + 4k, multiple of 16).
void launch_compose_keys(KeyColumns cols, uint32_t n, uint8_t *cells, uint32_t width,
                         hipStream_t stream);
// code[r] = kNoCode for the rows with some column code kNoCode.
void launch_mask_keys(KeyColumns cols, uint32_t n, uint32_t *code, hipStream_t stream);
// Unpacks n_rows rows of row_bytes bytes (the reference's row-major table on
// the device) into the columns' NUL-padded cells: each cell is copied up to
// its first NUL or `size` bytes, the rest of its width zeroed.  Work-groups
// stage rows_per_group rows at a time in LDS (unpack_lds_bytes of their
// span, at most kUnpackLds).
constexpr uint32_t kUnpackLds = 64 * 1024;
inline size_t unpack_lds_bytes(size_t span, size_t cells) {
  const size_t rows_room = ((span + 31) & ~size_t(31)) + 32;
  return rows_room + ((span + 31) / 32 + 1) * 4 + cells * 2;
}
// The column parameters of a table of at most kUnpackArgCols columns travel
// as kernel arguments (no copy ahead of the kernel); those of a wider table
// are read from device memory (cols_dev).
constexpr uint32_t kUnpackArgCols = 24;
struct UnpackColumns {
  UnpackColumn c[kUnpackArgCols];
};
void launch_unpack_rows(const uint8_t *rows, uint64_t row_bytes, uint32_t n_rows,
                        const UnpackColumns &cols, const UnpackColumn *cols_dev, uint32_t n_cols,
                        uint32_t rows_per_group, int n_cu, hipStream_t stream);
void launch_build_sig(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint16_t *len,
                      SigRow *sig, hipStream_t stream);
void launch_build_planes(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint32_t nwc,
                         uint32_t *planes, hipStream_t stream);
void launch_build_hist(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint32_t *hist,
                       hipStream_t stream);
// The three above in one pass over the cells (what the runtime uses; the
// separate kernels stay for the tests that compare them); hist may be null
// (no refine rows).
void launch_build_col(const uint8_t *bytes, uint32_t width, int64_t n_rows, uint32_t nwc,
                      uint16_t *len, SigRow *sig, uint32_t *hist, uint32_t *planes,
                      hipStream_t stream);
void launch_iota(uint32_t *out, uint32_t n, hipStream_t stream);
void launch_fill_u64(unsigned long long *p, unsigned long long n, unsigned long long v,
                     hipStream_t stream);

// Fills an allowance table (JRDesc::allow_tab) for coefficients c0, c3.
void launch_allow_table(float c0, float c3, uint32_t *tab, hipStream_t stream);

// Join order of a table for one rule shape: keys (code << 8) | min(length,
// 255) (code or length may be absent: null pointer), sorted with the row ids;
// then the runs of equal code (key >> 8): flags, exclusive scan, run arrays.
// A code not below code_bound sets *bad (the row then counts as code 0).
void launch_layout_keys_fb(const uint32_t *code, const SigRow *sig, const uint16_t *len,
                           uint32_t n, uint64_t code_bound, uint32_t *bad, uint64_t *keys,
                           hipStream_t stream);
void launch_layout_keys(const uint32_t *code, const uint16_t *len, uint32_t n,
                        uint64_t code_bound, uint32_t *bad, uint64_t *keys, hipStream_t stream);
void launch_gather_sig(const uint32_t *perm, uint32_t n, const SigRow *sig, SigRow *jsig,
                       hipStream_t stream);
// The whole layout of a table whose join keys take at most kLayoutMaxCodes
// codes, by counting sorts: join order (perm, jsig) by (code, length), the
// first-byte order (perm_fb, jsig_fb; with a similarity column) by (code,
// first byte), the runs of equal code and their count, in two launches.
// Layouts built one after another on one stream share the counters:
//   - up to kLayoutLdsCodes codes: hist (2 * kLayoutLdsCodes * 256
//     counters) must be zero; it is left holding this layout's counts, and
//     hist_clear (clear_words counters: the other of two such buffers, used
//     by the previous layout of this kind) is cleared, so consecutive
//     layouts alternate between the two;
//   - more codes: hist (2 * kLayoutMaxCodes * 256), *bad and *done must be
//     zero and are left zero; off holds 2 * n_codes * 256 entries.
// The layout's own flag *bad_out is written (1 when a code was not below
// n_codes).  rank holds 2 * n entries.  Within a bucket rows come in no
// particular order.
constexpr uint32_t kLayoutMaxCodes = 64;
constexpr uint32_t kLayoutLdsCodes = 28;  // counts and offsets in LDS (56 KiB) up to this many codes
struct LayoutSmall {
  const uint32_t *code;  // null: a rule without equalities (one code, 0)
  uint32_t *bad;         // more codes: set when a code is not below n_codes (row counts as code 0)
  uint32_t *bad_out;     // the layout's flag
  uint32_t *done;        // more codes: work-groups of the counting launch done
  const uint16_t *len;   // null: no similarity column
  const SigRow *sig;     // null: no similarity column
  uint32_t n, n_codes;
  uint32_t *hist, *rank, *off;
  uint32_t *hist_clear;  // up to kLayoutLdsCodes codes: counters cleared by this build
  uint32_t clear_words;
  uint32_t *perm, *perm_fb;
  SigRow *jsig, *jsig_fb;
  uint32_t *run_code, *run_off, *n_runs;
  uint32_t *n_runs_host;  // the run count also written here (pinned, device address)
};
void launch_layout_small(const LayoutSmall &a, hipStream_t stream);
void launch_run_flags(const uint64_t *sorted_keys, uint32_t n, uint32_t *flags,
                      hipStream_t stream);
// run_code / run_off of every run start (flags, run_id = exclusive scan of
// flags) and run_off[n_runs] = n, with n_runs = run_id[n-1] + flags[n-1]
// (also written to n_runs_host, pinned host memory at its device address).
void launch_layout_runs(const uint64_t *sorted_keys, const uint32_t *flags,
                        const uint32_t *run_id, uint32_t n, uint32_t *run_code,
                        uint32_t *run_off, uint32_t *n_runs, uint32_t *n_runs_host,
                        hipStream_t stream);
// Groups of every JR of a batch (one per left run), with their work.
// The first launch of a batch whose n + 1 groups fit one work-group of
// 1024 (not captured into a graph): copies desc_bytes of descriptors from
// pinned host memory (hdesc_dev: its device address; both 16-byte aligned,
// rounded up to 16 bytes), clears cnt_words counters and set_words dedupe
// slots, joins and scans the work items, in place of the copy, the fills,
// launch_batch_join and the scan.
constexpr uint32_t kPrologueGroups = 1023;
void launch_batch_prologue(BatchDev b, GroupArrays g, const void *hdesc_dev, void *ddesc,
                           size_t desc_bytes, unsigned long long *cnt, size_t cnt_words,
                           uint64_t set_words, hipStream_t stream);
void launch_batch_join(BatchDev b, GroupArrays g, hipStream_t stream);

// Filter variants: JRs without a similarity predicate; with one and an
// allowance table; with one and allowances computed in the kernel.
enum : int { kFilterEq = 0, kFilterSimTable = 1, kFilterSim = 2 };

// Stage 0 over groups [g_begin, g_end) (JRs of one filter variant): bound
// filter of the first similarity predicate (or every pair); kept pairs are
// appended to `out` tagged with their JR.
void launch_filter(BatchDev b, GroupArrays g, uint32_t g_begin, uint32_t g_end, int variant,
                   PairList out, FilterShape shape, int n_cu, hipStream_t stream);

// Second bound on the filter's survivors (`in`): the filter's JW bound for
// the rule's first similarity predicate with 128 byte classes instead of 32
// (for JRs with `refine` set; others pass); the pairs it does not reject go
// to `out`.
void launch_refine(BatchDev b, PairList in, PairList out, int n_cu, hipStream_t stream);

// Exact pass on list `in`: similarity predicates `stage` and `stage` + 1
// (1-based; the second where the pair's rule has it, evaluated only for
// pairs passing the first).  JRs without similarity predicates pass their
// pairs to the output directly; pairs passing their last predicate go to
// `final_` (deduplicated per job where needed), the others to `next`.
// Pairs with a shorter string above 256 bytes in either predicate run one
// wave per pair (second kernel, via `defer`; not launched when `may_defer`
// is false, i.e. no column pair can hold such a pair).  With
// `direct_pairs` set (a batch of one job) the final pairs go there as
// (left, right) at their final-list positions instead, and audit entries
// go to `direct_audit`: the job's output needs no scatter.
void launch_exact(BatchDev b, PairList in, int stage, PairList next, PairList final_,
                  PairList defer, AuditList audit, bool count_filtered, bool may_defer,
                  uint2 *direct_pairs, hb_audit_entry *direct_audit, int n_cu,
                  hipStream_t stream);

// Copies the final list and the audit list into per-job contiguous ranges:
// job j's pairs start at out_off[j] (exclusive scan of its match count).
void launch_scatter(BatchDev b, PairList final_, const uint64_t *out_off,
                    unsigned long long *fill, uint32_t *out_pairs, unsigned long long out_cap,
                    AuditList audit, const uint64_t *audit_off, unsigned long long *audit_fill,
                    hb_audit_entry *out_audit, unsigned long long audit_cap, int n_cu,
                    hipStream_t stream);

// Jaro-Winkler from counts, one thread per (m, S, T, p) tuple (tests).
void launch_jw_from_counts(const int32_t *mstp, int64_t n, double *out, double *out_fused,
                           hipStream_t stream);

// hb_sort.hip: rocPRIM wrappers.  temp may be null to query temp_bytes.
hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const uint64_t *keys_in,
                          uint64_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                          uint32_t n, int end_bit, hipStream_t stream);
hipError_t exscan_u32(void *temp, size_t &temp_bytes, const uint32_t *in, uint32_t *out,
                      uint32_t n, hipStream_t stream);
hipError_t exscan_u64(void *temp, size_t &temp_bytes, const uint64_t *in, uint64_t *out,
                      uint32_t n, hipStream_t stream);

}  // namespace hb
