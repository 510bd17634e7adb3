// C ABI runtime: contexts, device tables, plans and blocking jobs.
//
// Replaces the reference's submit/complete path: HostProducer::AsyncSubmit
// (core/components/host_producer.h:109-215) and HostReducer::Run
// (core/components/host_reducer.h:32-68).  Differences by design:
//   - device buffers come from a per-context cache, not a cudaMalloc per call
//     that is never freed;
//   - tables are copied host-to-device once (not twice) and may stay resident;
//   - results are index pairs with no 65,536 cap: a buffer that overflows is
//     grown and the job re-run;
//   - completion is an event per job, queried or waited on, not a busy poll.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "../../include/hyperblocker_amd.h"
#include "hb_internal.h"

namespace hbrt {

thread_local std::string g_last_error;

hb_status fail(hb_status st, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

#define HIP_OK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? HB_ERR_OOM : HB_ERR_HIP, "%s: %s (%s:%d)", \
                  #expr, hipGetErrorString(e_), __FILE__, __LINE__);                      \
  } while (0)

#define HB_TRY(expr)             \
  do {                           \
    hb_status s_ = (expr);       \
    if (s_ != HB_OK) return s_;  \
  } while (0)

// Device memory cache: blocks are kept per exact (rounded) size and reused.
class DevicePool {
 public:
  ~DevicePool() { release_all(); }
  hb_status alloc(size_t bytes, void **out) {
    bytes = round(bytes);
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = free_.find(bytes);
      if (it != free_.end() && !it->second.empty()) {
        *out = it->second.back();
        it->second.pop_back();
        return HB_OK;
      }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      // Give cached blocks back to the driver and retry once.
      release_all();
      e = hipMalloc(&p, bytes);
      if (e != hipSuccess)
        return fail(HB_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    *out = p;
    return HB_OK;
  }
  void free(void *p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_[round(bytes)].push_back(p);
  }
  void release_all() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &kv : free_)
      for (void *p : kv.second) (void)hipFree(p);
    free_.clear();
  }

 private:
  static size_t round(size_t b) {
    if (b < 4096) return 4096;
    size_t p = 4096;
    while (p < b) p <<= 1;
    return p;
  }
  std::mutex mu_;
  std::unordered_map<size_t, std::vector<void *>> free_;
};

// Pinned host memory cache (hipHostMalloc, mapped into the device's address
// space), reused across jobs: counters and job outputs live here, so results
// need no copy after the job and no allocation per job.
class HostPool {
 public:
  ~HostPool() { release_all(); }
  hb_status alloc(size_t bytes, void **host, void **dev) {
    bytes = round(bytes);
    void *p = nullptr;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = free_.find(bytes);
      if (it != free_.end() && !it->second.empty()) {
        p = it->second.back();
        it->second.pop_back();
      }
    }
    if (!p) {
      hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocMapped);
      if (e != hipSuccess) {
        release_all();
        e = hipHostMalloc(&p, bytes, hipHostMallocMapped);
        if (e != hipSuccess)
          return fail(HB_ERR_OOM, "hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
      }
    }
    *host = p;
    if (dev) {
      hipError_t e = hipHostGetDevicePointer(dev, p, 0);
      if (e != hipSuccess) {
        free(p, bytes);
        return fail(HB_ERR_HIP, "hipHostGetDevicePointer: %s", hipGetErrorString(e));
      }
    }
    return HB_OK;
  }
  void free(void *p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_[round(bytes)].push_back(p);
  }
  void release_all() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &kv : free_)
      for (void *p : kv.second) (void)hipHostFree(p);
    free_.clear();
  }

 private:
  static size_t round(size_t b) {
    size_t p = 256;
    while (p < b) p <<= 1;
    return p;
  }
  std::mutex mu_;
  std::unordered_map<size_t, std::vector<void *>> free_;
};

// Timing events, reused across jobs.
class EventPool {
 public:
  ~EventPool() {
    for (hipEvent_t e : free_) (void)hipEventDestroy(e);
  }
  hb_status get(hipEvent_t *e) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (!free_.empty()) {
        *e = free_.back();
        free_.pop_back();
        return HB_OK;
      }
    }
    HIP_OK(hipEventCreate(e));
    return HB_OK;
  }
  void put(hipEvent_t e) {
    if (!e) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_.push_back(e);
  }

 private:
  std::mutex mu_;
  std::vector<hipEvent_t> free_;
};

// Byte-string -> code dictionary shared by every table of a context, so that
// codes from any two columns compare equal exactly when the strings do.
class Dictionary {
 public:
  size_t size() const { return map_.size(); }
  uint32_t code(std::string_view s) {
    auto it = map_.find(s);
    if (it != map_.end()) return it->second;
    char *dst = store(s);
    uint32_t c = (uint32_t)map_.size();
    map_.emplace(std::string_view(dst, s.size()), c);
    return c;
  }
  std::mutex mu;

 private:
  char *store(std::string_view s) {
    if (blocks_.empty() || used_ + s.size() > kBlock) {
      blocks_.emplace_back(new char[std::max(kBlock, s.size())]);
      used_ = 0;
    }
    char *dst = blocks_.back().get() + used_;
    if (!s.empty()) std::memcpy(dst, s.data(), s.size());
    used_ += s.size();
    return dst;
  }
  static constexpr size_t kBlock = 1 << 20;
  std::unordered_map<std::string_view, uint32_t> map_;
  std::vector<std::unique_ptr<char[]>> blocks_;
  size_t used_ = 0;
};

}  // namespace hbrt

using namespace hbrt;

struct Job;

struct hb_ctx {
  int device = 0;
  int n_cu = 256;
  hb::FilterShape shape{2, 256, 16};  // filter launch shape; HB_FILTER_SHAPE overrides
  hipStream_t copy_stream = nullptr;
  std::vector<hipStream_t> streams;
  std::atomic<uint64_t> next_stream{0};
  DevicePool pool;
  HostPool hpool;
  EventPool events;
  Dictionary dict;
  std::mutex mu;  // jobs map
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
};

namespace {

// A lazily built per-column device structure, ready once `ev` completes.
struct Lazy {
  bool built = false;
  hipEvent_t ev = nullptr;
};

struct Column {
  uint32_t width = 16;
  uint8_t *d_bytes = nullptr;
  char *h_bytes = nullptr;  // pinned staging, kept for dictionary coding
  // similarity predicates
  Lazy sig;
  uint16_t *d_len = nullptr;
  hb::SigRow *d_sig = nullptr;
  uint32_t *d_planes = nullptr;  // bit-planes of the first 32*nwc bytes
  uint32_t nwc = 0;
  // equality predicates
  Lazy code;
  uint32_t *d_code = nullptr;
  uint32_t *h_code = nullptr;  // pinned
};

// The rows of a table in join order for one rule shape: sorted by the code
// of its first equality column, then by the length of its first similarity
// column (either may be absent, -1).  Kept for the table's lifetime.
struct Layout {
  Lazy lz;
  uint64_t *d_keys = nullptr;     // sorted (code << 8) | min(length, 255)
  uint32_t *d_perm = nullptr;     // row id of each join-order position
  hb::SigRow *d_jsig = nullptr;   // similarity signatures in join order
};

}  // namespace

struct hb_table {
  hb_ctx *ctx = nullptr;
  int64_t n_rows = 0;
  int32_t n_cols = 0;
  std::vector<Column> cols;
  std::map<std::pair<int, int>, Layout> layouts;  // (equality column, similarity column)
  hipEvent_t ready = nullptr;
  std::mutex mu;  // guards lazy builds
  std::atomic<int> users{0};
};

struct PlanRule {
  std::vector<std::pair<int32_t, int32_t>> eq;
  struct Sim {
    int32_t cl, cr, slot;
    float thr;
  };
  std::vector<Sim> sims;
};

struct hb_plan {
  int32_t n_rules = 0;
  std::vector<int32_t> pred_index;
  std::vector<char> pred_type;
  std::vector<float> pred_threshold;
  std::vector<PlanRule> rules;
  int32_t max_col_l = -1, max_col_r = -1;
};

namespace {

enum class JobState { kRunning, kDone, kFailed };

// Counter slots of a job (device, unsigned long long).
enum : int {
  kCntOverflow = 0,
  kCntJw = 1,
  kCntAudit = 2,
  kCntJoined = 3,
  kCntOut = 4,
  kCntFiltered = 5,
  kCntStage0 = 8,
};

}  // namespace

struct Job {
  uint64_t id = 0;
  hb_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  const hb_plan *plan = nullptr;
  hb_table *L = nullptr, *R = nullptr;
  std::mutex mu;
  JobState state = JobState::kRunning;
  hb_status error = HB_OK;
  std::string error_text;
  // capacities (pairs / entries)
  unsigned long long list_cap = 0, out_cap = 0, audit_cap = 0, set_cap = 0;
  int retries = 0;
  hb::FilterShape shape{2, 256, 16};  // the context's filter shape at submit
  // device memory
  struct Alloc {
    void *p;
    size_t n;
  };
  std::vector<Alloc> allocs;
  std::vector<Alloc> host_allocs;       // pinned, from the context's host pool
  unsigned long long *d_cnt = nullptr;
  unsigned long long *h_cnt = nullptr;  // pinned copy of the counters
  int n_counters = 0;
  std::vector<int> filter_counters;  // per rule: counter of stage 0, or -1
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  std::vector<hipEvent_t> marks;  // per rule: join, filter, exact stage boundaries
  size_t n_marks = 0;
  // outputs: pinned host memory written by the kernels through its device
  // address (d_*), read by the caller at the host address (h_*)
  uint32_t *d_out = nullptr, *h_out = nullptr;
  hb_audit_entry *d_audit = nullptr, *h_audit = nullptr;
  hb_job_stats stats{};
};

namespace {

hb_status mark(Job *j, hipStream_t st) {
  if (j->n_marks == j->marks.size()) {
    hipEvent_t e;
    HB_TRY(j->ctx->events.get(&e));
    j->marks.push_back(e);
  }
  HIP_OK(hipEventRecord(j->marks[j->n_marks++], st));
  return HB_OK;
}

hb_status job_alloc(Job *j, size_t bytes, void **out) {
  HB_TRY(j->ctx->pool.alloc(bytes, out));
  j->allocs.push_back({*out, bytes});
  return HB_OK;
}

hb_status job_alloc_host(Job *j, size_t bytes, void **host, void **dev) {
  HB_TRY(j->ctx->hpool.alloc(bytes, host, dev));
  j->host_allocs.push_back({*host, bytes});
  return HB_OK;
}

// Returns the job's device and pinned output memory to the context's pools
// (the job's stream must be idle).
void job_free_device(Job *j) {
  for (auto &a : j->allocs) j->ctx->pool.free(a.p, a.n);
  j->allocs.clear();
  for (auto &a : j->host_allocs) j->ctx->hpool.free(a.p, a.n);
  j->host_allocs.clear();
}

// Returns everything a finished job holds to the context's pools.
void job_free_all(Job *j) {
  job_free_device(j);
  if (j->h_cnt) j->ctx->hpool.free(j->h_cnt, (size_t)j->n_counters * sizeof(unsigned long long));
  j->h_cnt = nullptr;
  j->ctx->events.put(j->ev_start);
  j->ctx->events.put(j->ev_end);
  for (hipEvent_t e : j->marks) j->ctx->events.put(e);
  j->ev_start = j->ev_end = nullptr;
  j->marks.clear();
}

hb_status set_device(hb_ctx *ctx) {
  HIP_OK(hipSetDevice(ctx->device));
  return HB_OK;
}

// Waits (on `stream`) for a lazy structure built by another stream.
hb_status wait_lazy(const Lazy &lz, hipStream_t stream) {
  if (lz.ev) HIP_OK(hipStreamWaitEvent(stream, lz.ev, 0));
  return HB_OK;
}

hb_status record_lazy(Lazy &lz, hipStream_t stream) {
  if (!lz.ev) HIP_OK(hipEventCreateWithFlags(&lz.ev, hipEventDisableTiming));
  HIP_OK(hipEventRecord(lz.ev, stream));
  lz.built = true;
  return HB_OK;
}

hb_status ensure_sig(hb_table *t, int c, hipStream_t stream) {
  Column &col = t->cols[c];
  std::lock_guard<std::mutex> lk(t->mu);
  if (col.sig.built) return wait_lazy(col.sig, stream);
  const size_t n = (size_t)std::max<int64_t>(t->n_rows, 1);
  HIP_OK(hipMalloc(&col.d_len, n * sizeof(uint16_t)));
  HIP_OK(hipMalloc(&col.d_sig, n * sizeof(hb::SigRow)));
  col.nwc = std::min<uint32_t>((col.width + 31) / 32, hb::kMaxPlaneWords);
  HIP_OK(hipMalloc(&col.d_planes, n * col.nwc * 32));
  hb::launch_build_sig(col.d_bytes, col.width, t->n_rows, col.d_len, col.d_sig, stream);
  hb::launch_build_planes(col.d_bytes, col.width, t->n_rows, col.nwc, col.d_planes, stream);
  HIP_OK(hipGetLastError());
  return record_lazy(col.sig, stream);
}

hb_status ensure_code(hb_table *t, int c, hipStream_t stream) {
  Column &col = t->cols[c];
  std::lock_guard<std::mutex> lk(t->mu);
  if (col.code.built) return wait_lazy(col.code, stream);
  const size_t n = (size_t)t->n_rows;
  HIP_OK(hipHostMalloc((void **)&col.h_code, std::max<size_t>(n, 1) * sizeof(uint32_t),
                       hipHostMallocDefault));
  {
    Dictionary &d = t->ctx->dict;
    std::lock_guard<std::mutex> dl(d.mu);
    for (size_t r = 0; r < n; r++) {
      const char *s = col.h_bytes + r * col.width;
      col.h_code[r] = d.code(std::string_view(s, strnlen(s, col.width)));
    }
  }
  HIP_OK(hipMalloc(&col.d_code, std::max<size_t>(n, 1) * sizeof(uint32_t)));
  if (n) HIP_OK(hipMemcpyAsync(col.d_code, col.h_code, n * sizeof(uint32_t),
                               hipMemcpyHostToDevice, stream));
  return record_lazy(col.code, stream);
}

// Builds (once per table and rule shape) the join-order layout: keys from the
// equality column's dictionary codes and the similarity column's lengths,
// radix-sorted with the row ids, and the similarity signatures gathered into
// that order.
hb_status ensure_layout(hb_table *t, int eq, int sim, hipStream_t stream, const Layout **out) {
  if (eq >= 0) HB_TRY(ensure_code(t, eq, stream));
  if (sim >= 0) HB_TRY(ensure_sig(t, sim, stream));
  uint64_t n_codes = 1;
  if (eq >= 0) {
    Dictionary &d = t->ctx->dict;
    std::lock_guard<std::mutex> dl(d.mu);
    n_codes = std::max<uint64_t>(d.size(), 1);
  }
  std::lock_guard<std::mutex> lk(t->mu);
  Layout &ly = t->layouts[{eq, sim}];
  *out = &ly;
  if (ly.lz.built) return wait_lazy(ly.lz, stream);
  const uint32_t n = (uint32_t)t->n_rows;
  const size_t n1 = std::max<size_t>(n, 1);
  HIP_OK(hipMalloc(&ly.d_keys, n1 * sizeof(uint64_t)));
  HIP_OK(hipMalloc(&ly.d_perm, n1 * sizeof(uint32_t)));
  if (sim >= 0) HIP_OK(hipMalloc(&ly.d_jsig, n1 * sizeof(hb::SigRow)));
  if (n) {
    uint64_t *keys_in = nullptr;
    uint32_t *iota = nullptr;
    HIP_OK(hipMalloc(&keys_in, n1 * sizeof(uint64_t)));
    HIP_OK(hipMalloc(&iota, n1 * sizeof(uint32_t)));
    hb::launch_layout_keys(eq >= 0 ? t->cols[eq].d_code : nullptr,
                           sim >= 0 ? t->cols[sim].d_len : nullptr, n, keys_in, stream);
    hb::launch_iota(iota, n, stream);
    int end_bit = 8;
    while (end_bit < 64 && (1ull << (end_bit - 8)) < n_codes) end_bit++;
    size_t tb = 0;
    HIP_OK(hb::sort_pairs_u64(nullptr, tb, keys_in, ly.d_keys, iota, ly.d_perm, n, end_bit,
                              stream));
    void *tmp = nullptr;
    HIP_OK(hipMalloc(&tmp, std::max<size_t>(tb, 16)));
    HIP_OK(hb::sort_pairs_u64(tmp, tb, keys_in, ly.d_keys, iota, ly.d_perm, n, end_bit, stream));
    if (sim >= 0) hb::launch_gather_sig(ly.d_perm, n, t->cols[sim].d_sig, ly.d_jsig, stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(stream));  // one-off per layout; frees below need it
    HIP_OK(hipFree(tmp));
    HIP_OK(hipFree(keys_in));
    HIP_OK(hipFree(iota));
  }
  return record_lazy(ly.lz, stream);
}

float filter_coef(double thr, int p) {
  const double jn = (thr - 2e-6 - 0.1 * p) / (1.0 - 0.1 * p);
  double c = 3.0 * jn - 1.0;
  c -= std::fabs(c) * 1e-6 + 1e-9;  // strictly below the exact boundary
  float cf = (float)c;
  if ((double)cf > c) cf = std::nextafter(cf, -INFINITY);
  return cf;
}

bool valid_filter_shape(int rows_per_lane, int tile_b, int groups_per_cu) {
  return (rows_per_lane == 1 || rows_per_lane == 2 || rows_per_lane == 4) && tile_b >= 256 &&
         tile_b <= 65536 && tile_b % 256 == 0 && groups_per_cu >= 1 && groups_per_cu <= 64;
}

hb_status enqueue_job(Job *j) {
  hb_ctx *ctx = j->ctx;
  const hb_plan *plan = j->plan;
  hb_table *L = j->L, *R = j->R;
  hipStream_t st = j->stream;
  const uint32_t n_l = (uint32_t)L->n_rows, n_r = (uint32_t)R->n_rows;

  HIP_OK(hipStreamWaitEvent(st, L->ready, 0));
  if (R != L) HIP_OK(hipStreamWaitEvent(st, R->ready, 0));
  // Lazy per-column structures and join layouts needed by the plan.
  std::vector<std::pair<const Layout *, const Layout *>> layouts;
  for (const PlanRule &rule : plan->rules) {
    for (size_t k = 1; k < rule.eq.size(); k++) {
      HB_TRY(ensure_code(L, rule.eq[k].first, st));
      HB_TRY(ensure_code(R, rule.eq[k].second, st));
    }
    for (const auto &s : rule.sims) {
      HB_TRY(ensure_sig(L, s.cl, st));
      HB_TRY(ensure_sig(R, s.cr, st));
    }
    const int el = rule.eq.empty() ? -1 : rule.eq[0].first;
    const int er = rule.eq.empty() ? -1 : rule.eq[0].second;
    const int sl = rule.sims.empty() ? -1 : rule.sims[0].cl;
    const int sr = rule.sims.empty() ? -1 : rule.sims[0].cr;
    const Layout *ll = nullptr, *lr = nullptr;
    HB_TRY(ensure_layout(L, el, sl, st, &ll));
    HB_TRY(ensure_layout(R, er, sr, st, &lr));
    layouts.push_back({ll, lr});
  }

  // Job memory.
  void *p = nullptr;
  j->filter_counters.clear();
  HB_TRY(job_alloc(j, (size_t)j->n_counters * sizeof(unsigned long long), &p));
  j->d_cnt = (unsigned long long *)p;
  uint32_t *list[2];
  HB_TRY(job_alloc(j, j->list_cap * 8, &p));
  list[0] = (uint32_t *)p;
  HB_TRY(job_alloc(j, j->list_cap * 8, &p));
  list[1] = (uint32_t *)p;
  HB_TRY(job_alloc(j, j->list_cap * 8, &p));
  uint32_t *d_defer = (uint32_t *)p;
  void *hp = nullptr;
  HB_TRY(job_alloc_host(j, j->out_cap * 8, &hp, &p));
  j->h_out = (uint32_t *)hp;
  j->d_out = (uint32_t *)p;
  HB_TRY(job_alloc_host(j, j->audit_cap * sizeof(hb_audit_entry), &hp, &p));
  j->h_audit = (hb_audit_entry *)hp;
  j->d_audit = (hb_audit_entry *)p;
  unsigned long long *d_set = nullptr;
  const bool multi = plan->rules.size() > 1;
  if (multi) {
    HB_TRY(job_alloc(j, j->set_cap * 8, &p));
    d_set = (unsigned long long *)p;
  }
  // Join scratch sized for the larger of the left tables' row counts.
  const uint32_t cap = std::max<uint32_t>(n_l, 1);
  uint32_t *flags, *run_id;
  HB_TRY(job_alloc(j, (size_t)cap * 4, &p));
  flags = (uint32_t *)p;
  HB_TRY(job_alloc(j, (size_t)cap * 4, &p));
  run_id = (uint32_t *)p;
  // Extra equality codes in join order (the signatures come with the layout).
  uint32_t *jcode_l, *jcode_r;
  HB_TRY(job_alloc(j, (size_t)std::max<uint32_t>(n_l, 1) * 4 * hb::kMaxEq, &p));
  jcode_l = (uint32_t *)p;
  HB_TRY(job_alloc(j, (size_t)std::max<uint32_t>(n_r, 1) * 4 * hb::kMaxEq, &p));
  jcode_r = (uint32_t *)p;
  hb::GroupArrays g;
  g.cap = n_l;
  g.tile_a = 256u * (uint32_t)j->shape.rows_per_lane;
  g.tile_b = j->shape.tile_b;
  g.small_pairs_max = 16384;
  HB_TRY(job_alloc(j, 4 * sizeof(uint32_t), &p));
  g.n_runs = (uint32_t *)p;
  uint32_t *u32arr[4];
  for (auto &a : u32arr) {
    HB_TRY(job_alloc(j, ((size_t)cap + 2) * 4, &p));
    a = (uint32_t *)p;
  }
  g.l_off = u32arr[0];
  g.l_cnt = u32arr[1];
  g.r_off = u32arr[2];
  g.r_cnt = u32arr[3];
  uint64_t *u64arr[4];
  for (auto &a : u64arr) {
    HB_TRY(job_alloc(j, ((size_t)cap + 2) * 8, &p));
    a = (uint64_t *)p;
  }
  g.small_pairs = u64arr[0];
  g.pair_off = u64arr[1];
  g.items = u64arr[2];
  g.item_off = u64arr[3];
  g.joined = j->d_cnt + kCntJoined;
  size_t tb32 = 0, tb64 = 0;
  HIP_OK(hb::exscan_u32(nullptr, tb32, flags, run_id, cap, st));
  HIP_OK(hb::exscan_u64(nullptr, tb64, g.small_pairs, g.pair_off, std::max<uint32_t>(cap + 1, 2),
                        st));
  void *tmp = nullptr;
  const size_t tmp_bytes = std::max<size_t>(std::max(tb32, tb64), 16);
  HB_TRY(job_alloc(j, tmp_bytes, &tmp));

  j->n_marks = 0;
  HIP_OK(hipEventRecord(j->ev_start, st));
  hb::launch_fill_u64(j->d_cnt, (unsigned long long)j->n_counters, 0ull, st);
  if (multi) hb::launch_fill_u64(d_set, j->set_cap, ~0ull, st);

  hb::AuditOut audit{j->d_audit, j->d_cnt + kCntAudit, j->audit_cap};
  hb::StageOut job_out{j->d_out, j->d_cnt + kCntOut, j->out_cap, j->d_cnt + kCntOverflow};
  int stage = kCntStage0;

  for (size_t ri = 0; ri < plan->rules.size(); ri++) {
    const PlanRule &rule = plan->rules[ri];
    const Layout &ll = *layouts[ri].first, &lr = *layouts[ri].second;
    const uint32_t *perm_l = ll.d_perm, *perm_r = lr.d_perm;
    hb::EqCheck eq{};
    HB_TRY(mark(j, st));
    if (!rule.eq.empty()) {
      if (n_l) {
        hb::launch_run_flags(ll.d_keys, n_l, flags, st);
        size_t t = tmp_bytes;
        HIP_OK(hb::exscan_u32(tmp, t, flags, run_id, n_l, st));
      }
      hb::launch_run_info(ll.d_keys, flags, run_id, n_l, lr.d_keys, n_r, g, st);
      eq.n = (int32_t)rule.eq.size() - 1;
      for (int k = 1; k < (int)rule.eq.size(); k++) {
        eq.code_l[k - 1] = L->cols[rule.eq[k].first].d_code;
        eq.code_r[k - 1] = R->cols[rule.eq[k].second].d_code;
      }
    } else {
      hb::GroupArrays g1 = g;
      g1.cap = 1;
      hb::launch_single_group(n_l, n_r, g1, st);
    }
    hb::GroupArrays gu = g;
    if (rule.eq.empty()) gu.cap = 1;
    {
      size_t t = tmp_bytes;
      HIP_OK(hb::exscan_u64(tmp, t, gu.small_pairs, gu.pair_off, gu.cap + 1, st));
      t = tmp_bytes;
      HIP_OK(hb::exscan_u64(tmp, t, gu.items, gu.item_off, gu.cap + 1, st));
    }
    // Stage chain: stage 0 = equalities + bound filter, stage s >= 1 = exact
    // JW of similarity s-1 on the previous stage's list.  Stages alternate
    // between the two lists; the last stage of a single-rule plan writes the
    // job output directly.
    const int n_stages = 1 + (int)rule.sims.size();
    float coef[4] = {0, 0, 0, 0};
    if (!rule.sims.empty())
      for (int q = 0; q < 4; q++) coef[q] = filter_coef((double)rule.sims[0].thr, q);
    const hb::SigRow *jl = ll.d_jsig, *jr = lr.d_jsig;
    if (eq.n > 0) {
      hb::launch_gather_join(perm_l, n_l, nullptr, nullptr, eq, 0, jcode_l, st);
      hb::launch_gather_join(perm_r, n_r, nullptr, nullptr, eq, 1, jcode_r, st);
      HIP_OK(hipGetLastError());
    }
    hb::StageOut prev{};
    HB_TRY(mark(j, st));
    for (int s = 0; s < n_stages; s++) {
      hb::StageOut o;
      if (!multi && s == n_stages - 1) {
        o = job_out;
        if (s == 0) j->filter_counters.push_back(-1);
      } else {
        if (stage >= j->n_counters)
          return fail(HB_ERR_INVALID_ARG, "internal: stage counters exhausted");
        o = hb::StageOut{list[s & 1], j->d_cnt + stage, j->list_cap, j->d_cnt + kCntOverflow};
        if (s == 0) j->filter_counters.push_back(stage);
        stage++;
      }
      if (s == 0) {
        hb::launch_filter_groups(gu, perm_l, perm_r, jl, jr, jcode_l, jcode_r, eq.n,
                                 rule.sims.empty() ? 0 : 1, coef, o, j->shape, ctx->n_cu, st);
      } else {
        const auto &sim = rule.sims[s - 1];
        const Column &cl = L->cols[sim.cl];
        const Column &cr = R->cols[sim.cr];
        hb::DevColumn dl{cl.d_bytes, cl.d_len, cl.width, cl.d_planes, cl.nwc};
        hb::DevColumn dr{cr.d_bytes, cr.d_len, cr.width, cr.d_planes, cr.nwc};
        if (stage >= j->n_counters)
          return fail(HB_ERR_INVALID_ARG, "internal: stage counters exhausted");
        hb::StageOut defer{d_defer, j->d_cnt + stage, j->list_cap, j->d_cnt + kCntOverflow};
        stage++;
        hb::launch_exact(prev.pairs, prev.count, prev.cap, dl, dr, sim.thr, sim.slot, o, audit,
                         j->d_cnt + kCntJw, defer, ctx->n_cu, st);
      }
      HIP_OK(hipGetLastError());
      prev = o;
      if (s == 0) HB_TRY(mark(j, st));
    }
    HB_TRY(mark(j, st));
    if (multi) {
      hb::launch_merge_unique(prev.pairs, prev.count, prev.cap, d_set, j->set_cap, job_out,
                              ctx->n_cu, st);
      HIP_OK(hipGetLastError());
    }
  }
  HIP_OK(hipMemcpyAsync(j->h_cnt, j->d_cnt, (size_t)j->n_counters * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, st));
  HIP_OK(hipEventRecord(j->ev_end, st));
  return HB_OK;
}

// Called with j->mu held once ev_end has completed.
hb_status job_complete(Job *j) {
  const unsigned long long *c = j->h_cnt;
  if (c[kCntOverflow]) {
    unsigned long long need_list = 0;
    for (int s = kCntStage0; s < j->n_counters; s++) need_list = std::max(need_list, c[s]);
    auto grow = [](unsigned long long cur, unsigned long long need) {
      unsigned long long n = cur;
      while (n < need + need / 4 + 1024) n *= 2;
      return n;
    };
    j->list_cap = grow(j->list_cap, need_list);
    j->out_cap = grow(j->out_cap, c[kCntOut]);
    j->audit_cap = grow(j->audit_cap, c[kCntAudit]);
    unsigned long long sc = 1024;
    while (sc < 2 * j->out_cap) sc <<= 1;
    j->set_cap = sc;
    j->retries++;
    if (j->retries > 8) return fail(HB_ERR_CAPACITY, "job %llu: output keeps overflowing",
                                    (unsigned long long)j->id);
    HIP_OK(hipStreamSynchronize(j->stream));
    job_free_device(j);
    HB_TRY(enqueue_job(j));
    return HB_NOT_READY;
  }
  float ms = 0;
  (void)hipEventElapsedTime(&ms, j->ev_start, j->ev_end);
  hb_job_stats &s = j->stats;
  s.pairs = (unsigned long long)j->L->n_rows * (unsigned long long)j->R->n_rows;
  s.candidates = c[kCntJoined];
  unsigned long long filt = 0;
  for (size_t k = 0; k < j->filter_counters.size(); k++) {
    // stage 0 of a rule with similarity predicates always writes a list
    if (j->filter_counters[k] >= 0 && !j->plan->rules[k].sims.empty())
      filt += c[j->filter_counters[k]];
  }
  s.filtered = filt;
  s.jw_exact = c[kCntJw];
  s.matches = c[kCntOut];
  s.audit = c[kCntAudit];
  s.device_ms = ms;
  s.join_ms = s.filter_ms = s.exact_ms = 0;
  for (size_t k = 0; k + 4 <= j->n_marks; k += 4) {  // 4 marks per rule
    float a = 0, b = 0, c2 = 0;
    (void)hipEventElapsedTime(&a, j->marks[k], j->marks[k + 1]);
    (void)hipEventElapsedTime(&b, j->marks[k + 1], j->marks[k + 2]);
    (void)hipEventElapsedTime(&c2, j->marks[k + 2], j->marks[k + 3]);
    s.join_ms += a;
    s.filter_ms += b;
    s.exact_ms += c2;
  }
  s.retries = j->retries;
  s.n_rules = (int32_t)j->plan->rules.size();
  j->state = JobState::kDone;
  return HB_OK;
}

hb_status job_poll(Job *j, bool block) {
  std::lock_guard<std::mutex> lk(j->mu);
  for (;;) {
    if (j->state == JobState::kDone) return HB_OK;
    if (j->state == JobState::kFailed)
      return fail(j->error, "%s", j->error_text.c_str());
    HB_TRY(set_device(j->ctx));
    hipError_t e = block ? hipEventSynchronize(j->ev_end) : hipEventQuery(j->ev_end);
    if (e == hipErrorNotReady) return HB_NOT_READY;
    if (e != hipSuccess) {
      j->state = JobState::kFailed;
      j->error = HB_ERR_HIP;
      j->error_text = std::string("job failed on device: ") + hipGetErrorString(e);
      return fail(j->error, "%s", j->error_text.c_str());
    }
    hb_status st = job_complete(j);
    if (st == HB_NOT_READY) {
      if (!block) return HB_NOT_READY;
      continue;
    }
    if (st != HB_OK) {
      j->state = JobState::kFailed;
      j->error = st;
      j->error_text = g_last_error;
      return st;
    }
  }
}

Job *find_job(hb_ctx *ctx, uint64_t id) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->jobs.find(id);
  return it == ctx->jobs.end() ? nullptr : it->second.get();
}

hb_status check_cell_bytes(const hb_table_view *v) {
  if (!v || v->n_rows < 0 || v->n_cols <= 0 || !v->col_base || !v->col_stride)
    return fail(HB_ERR_INVALID_ARG, "table view: bad shape or null arrays");
  if (v->n_rows > 0xFFFFFFF0ll)
    return fail(HB_ERR_CAPACITY, "table view: %lld rows exceed the 32-bit row index",
                (long long)v->n_rows);
  for (int c = 0; c < v->n_cols; c++)
    if ((v->n_rows > 0 && !v->col_base[c]) || v->col_stride[c] <= 0)
      return fail(HB_ERR_INVALID_ARG, "table view: column %d has no data or stride <= 0", c);
  return HB_OK;
}

}  // namespace

extern "C" {

const char *hb_last_error(void) { return g_last_error.c_str(); }

const char *hb_version(void) { return "hyperblocker_amd 0.1 gfx950"; }

hb_status hb_ctx_create(int32_t device, hb_ctx **out) {
  if (!out) return fail(HB_ERR_INVALID_ARG, "hb_ctx_create: out is null");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(HB_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= n)
    return fail(HB_ERR_INVALID_ARG, "device %d out of range [0, %d)", device, n);
  HIP_OK(hipSetDevice(device));
  auto ctx = std::make_unique<hb_ctx>();
  ctx->device = device;
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, device));
  ctx->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (const char *env = std::getenv("HB_FILTER_SHAPE")) {
    int rpl = 0, tb = 0, gpc = 0;
    if (std::sscanf(env, "%d,%d,%d", &rpl, &tb, &gpc) == 3 && valid_filter_shape(rpl, tb, gpc))
      ctx->shape = hb::FilterShape{rpl, (uint32_t)tb, (unsigned)gpc};
  }
  HIP_OK(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
  for (int i = 0; i < 4; i++) {
    hipStream_t s;
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    ctx->streams.push_back(s);
  }
  *out = ctx.release();
  return HB_OK;
}

hb_status hb_ctx_set_filter_shape(hb_ctx *ctx, int32_t rows_per_lane, int32_t tile_b,
                                  int32_t groups_per_cu) {
  if (!ctx) return fail(HB_ERR_INVALID_ARG, "hb_ctx_set_filter_shape: null context");
  if (!valid_filter_shape(rows_per_lane, tile_b, groups_per_cu))
    return fail(HB_ERR_INVALID_ARG,
                "filter shape (%d, %d, %d): rows per lane must be 1, 2 or 4, tile_b a multiple "
                "of 256 up to 65536, groups per CU 1..64",
                rows_per_lane, tile_b, groups_per_cu);
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->shape = hb::FilterShape{rows_per_lane, (uint32_t)tile_b, (unsigned)groups_per_cu};
  return HB_OK;
}

hb_status hb_ctx_get_filter_shape(const hb_ctx *ctx, int32_t *rows_per_lane, int32_t *tile_b,
                                  int32_t *groups_per_cu) {
  if (!ctx) return fail(HB_ERR_INVALID_ARG, "hb_ctx_get_filter_shape: null context");
  if (rows_per_lane) *rows_per_lane = ctx->shape.rows_per_lane;
  if (tile_b) *tile_b = (int32_t)ctx->shape.tile_b;
  if (groups_per_cu) *groups_per_cu = (int32_t)ctx->shape.grid_per_cu;
  return HB_OK;
}

hb_status hb_ctx_synchronize(hb_ctx *ctx) {
  if (!ctx) return fail(HB_ERR_INVALID_ARG, "hb_ctx_synchronize: null context");
  HB_TRY(set_device(ctx));
  HIP_OK(hipDeviceSynchronize());
  return HB_OK;
}

hb_status hb_ctx_destroy(hb_ctx *ctx) {
  if (!ctx) return HB_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (auto &kv : ctx->jobs) job_free_all(kv.second.get());
    ctx->jobs.clear();
  }
  ctx->pool.release_all();
  ctx->hpool.release_all();
  for (hipStream_t s : ctx->streams) (void)hipStreamDestroy(s);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  delete ctx;
  return HB_OK;
}

hb_status hb_table_upload_rows(hb_ctx *ctx, const hb_table_view *v, const int64_t *rows,
                               int64_t n_rows, hb_table **out) {
  if (!ctx || !out) return fail(HB_ERR_INVALID_ARG, "hb_table_upload: null argument");
  *out = nullptr;
  HB_TRY(check_cell_bytes(v));
  if (rows) {
    if (n_rows < 0 || n_rows > 0xFFFFFFF0ll)
      return fail(HB_ERR_INVALID_ARG, "hb_table_upload_rows: bad row count %lld",
                  (long long)n_rows);
    for (int64_t i = 0; i < n_rows; i++)
      if (rows[i] < 0 || rows[i] >= v->n_rows)
        return fail(HB_ERR_INVALID_ARG, "hb_table_upload_rows: row %lld out of range",
                    (long long)rows[i]);
  } else {
    n_rows = v->n_rows;
  }
  HB_TRY(set_device(ctx));
  auto t = std::make_unique<hb_table>();
  t->ctx = ctx;
  t->n_rows = n_rows;
  t->n_cols = v->n_cols;
  t->cols.resize(v->n_cols);
  auto src_row = [&](int64_t i) { return rows ? rows[i] : i; };
  for (int c = 0; c < v->n_cols; c++) {
    Column &col = t->cols[c];
    const int64_t stride = v->col_stride[c];
    size_t maxlen = 0;
    for (int64_t i = 0; i < n_rows; i++)
      maxlen = std::max(maxlen, strnlen(v->col_base[c] + src_row(i) * stride, (size_t)stride));
    col.width = (uint32_t)((maxlen + 1 + 15) & ~size_t(15));
    const size_t bytes = std::max<size_t>((size_t)n_rows * col.width, 16);
    HIP_OK(hipHostMalloc((void **)&col.h_bytes, bytes, hipHostMallocDefault));
    std::memset(col.h_bytes, 0, bytes);
    for (int64_t i = 0; i < n_rows; i++) {
      const char *s = v->col_base[c] + src_row(i) * stride;
      std::memcpy(col.h_bytes + i * col.width, s, strnlen(s, (size_t)stride));
    }
    HIP_OK(hipMalloc(&col.d_bytes, bytes));
    HIP_OK(hipMemcpyAsync(col.d_bytes, col.h_bytes, bytes, hipMemcpyHostToDevice,
                          ctx->copy_stream));
  }
  HIP_OK(hipEventCreateWithFlags(&t->ready, hipEventDisableTiming));
  HIP_OK(hipEventRecord(t->ready, ctx->copy_stream));
  *out = t.release();
  return HB_OK;
}

hb_status hb_table_upload(hb_ctx *ctx, const hb_table_view *v, hb_table **out) {
  return hb_table_upload_rows(ctx, v, nullptr, 0, out);
}

hb_status hb_table_upload_serialized(hb_ctx *ctx, const hb_serialized_table *s, hb_table **out) {
  if (!s || (!s->data && s->n_rows) || !s->col_size || !s->col_offset)
    return fail(HB_ERR_INVALID_ARG, "hb_table_upload_serialized: null argument");
  std::vector<const char *> base(s->n_cols);
  std::vector<int64_t> stride(s->n_cols);
  for (size_t c = 0; c < s->n_cols; c++) {
    if (s->col_offset[c] + s->col_size[c] > s->aligned_tuple_size)
      return fail(HB_ERR_INVALID_ARG, "column %zu exceeds aligned_tuple_size", c);
    base[c] = s->data + s->col_offset[c];
    stride[c] = (int64_t)s->aligned_tuple_size;
  }
  // Each cell is capped at its col_size: copy through a view whose stride is
  // the row size but whose cells stop at col_size bytes.
  std::vector<std::vector<char>> cols(s->n_cols);
  std::vector<const char *> cb(s->n_cols);
  std::vector<int64_t> cs(s->n_cols);
  for (size_t c = 0; c < s->n_cols; c++) {
    const size_t w = s->col_size[c] + 1;
    cols[c].assign(std::max<size_t>(s->n_rows * w, 1), 0);
    for (size_t r = 0; r < s->n_rows; r++)
      std::memcpy(cols[c].data() + r * w, base[c] + r * s->aligned_tuple_size, s->col_size[c]);
    cb[c] = cols[c].data();
    cs[c] = (int64_t)w;
  }
  hb_table_view v{(int64_t)s->n_rows, (int32_t)s->n_cols, cb.data(), cs.data()};
  return hb_table_upload(ctx, &v, out);
}

hb_status hb_table_info(const hb_table *t, int64_t *n_rows, int32_t *n_cols) {
  if (!t) return fail(HB_ERR_INVALID_ARG, "hb_table_info: null table");
  if (n_rows) *n_rows = t->n_rows;
  if (n_cols) *n_cols = t->n_cols;
  return HB_OK;
}

hb_status hb_table_release(hb_ctx *ctx, hb_table *t) {
  if (!ctx || !t) return fail(HB_ERR_INVALID_ARG, "hb_table_release: null argument");
  HB_TRY(set_device(ctx));
  // Jobs hold no table memory of their own, so every stream that may still
  // read this table must drain first.
  for (hipStream_t s : ctx->streams) HIP_OK(hipStreamSynchronize(s));
  HIP_OK(hipStreamSynchronize(ctx->copy_stream));
  for (Column &c : t->cols) {
    (void)hipFree(c.d_bytes);
    if (c.h_bytes) (void)hipHostFree(c.h_bytes);
    (void)hipFree(c.d_len);
    (void)hipFree(c.d_sig);
    (void)hipFree(c.d_planes);
    (void)hipFree(c.d_code);
    if (c.h_code) (void)hipHostFree(c.h_code);
    for (Lazy *lz : {&c.sig, &c.code})
      if (lz->ev) (void)hipEventDestroy(lz->ev);
  }
  for (auto &kv : t->layouts) {
    (void)hipFree(kv.second.d_keys);
    (void)hipFree(kv.second.d_perm);
    (void)hipFree(kv.second.d_jsig);
    if (kv.second.lz.ev) (void)hipEventDestroy(kv.second.lz.ev);
  }
  if (t->ready) (void)hipEventDestroy(t->ready);
  delete t;
  return HB_OK;
}

hb_status hb_plan_create(hb_ctx *ctx, const hb_serialized_plan *sp, hb_plan **out) {
  if (!out || !sp) return fail(HB_ERR_INVALID_ARG, "hb_plan_create: null argument");
  (void)ctx;
  *out = nullptr;
  if (sp->length <= 0 || sp->n_rules <= 0 || !sp->pred_index || !sp->pred_type ||
      !sp->pred_threshold)
    return fail(HB_ERR_PLAN, "plan: empty (n_rules=%d, length=%d)", sp->n_rules, sp->length);
  auto plan = std::make_unique<hb_plan>();
  plan->pred_index.assign(sp->pred_index, sp->pred_index + sp->length);
  plan->pred_type.assign(sp->pred_type, sp->pred_type + sp->length);
  plan->pred_threshold.assign(sp->pred_threshold, sp->pred_threshold + sp->length);
  if (sp->pred_index[sp->length - 1] != HB_CHECK_POINT)
    return fail(HB_ERR_PLAN, "plan: last slot is not a checkpoint");
  int i = 0;
  while (i < sp->length) {
    if (sp->pred_index[i] == HB_CHECK_POINT)
      return fail(HB_ERR_PLAN, "plan: rule %zu is empty (slot %d)", plan->rules.size(), i);
    PlanRule rule;
    while (i < sp->length && sp->pred_index[i] != HB_CHECK_POINT) {
      if (i + 1 >= sp->length || sp->pred_index[i + 1] == HB_CHECK_POINT)
        return fail(HB_ERR_PLAN, "plan: slot %d has no partner slot", i);
      const char t0 = sp->pred_type[i], t1 = sp->pred_type[i + 1];
      if (t0 != HB_EQUALITIES && t0 != HB_SIM)
        return fail(HB_ERR_PLAN, "plan: slot %d has predicate type 0x%02x", i, (unsigned char)t0);
      if (t0 != t1) return fail(HB_ERR_PLAN, "plan: slots %d/%d differ in type", i, i + 1);
      const int32_t cl = sp->pred_index[i], cr = sp->pred_index[i + 1];
      if (cl < 0 || cr < 0) return fail(HB_ERR_PLAN, "plan: negative column at slot %d", i);
      plan->max_col_l = std::max(plan->max_col_l, cl);
      plan->max_col_r = std::max(plan->max_col_r, cr);
      if (t0 == HB_EQUALITIES) {
        rule.eq.emplace_back(cl, cr);
      } else {
        rule.sims.push_back({cl, cr, i, sp->pred_threshold[i]});
      }
      i += 2;
    }
    i++;  // checkpoint
    if ((int)rule.eq.size() > 1 + hb::kMaxEq)
      return fail(HB_ERR_CAPACITY, "plan: rule %zu has %zu equalities (at most %d)",
                  plan->rules.size(), rule.eq.size(), 1 + hb::kMaxEq);
    if (rule.sims.size() > 16)
      return fail(HB_ERR_CAPACITY, "plan: rule %zu has %zu similarity predicates (at most 16)",
                  plan->rules.size(), rule.sims.size());
    plan->rules.push_back(std::move(rule));
  }
  if ((int)plan->rules.size() != sp->n_rules)
    return fail(HB_ERR_PLAN, "plan: n_rules=%d but %zu rules found", sp->n_rules,
                plan->rules.size());
  plan->n_rules = sp->n_rules;
  *out = plan.release();
  return HB_OK;
}

hb_status hb_plan_release(hb_ctx *ctx, hb_plan *plan) {
  (void)ctx;
  delete plan;
  return HB_OK;
}

hb_status hb_job_submit(hb_ctx *ctx, const hb_plan *plan, const hb_table *l, const hb_table *r,
                        uint64_t job_id) {
  if (!ctx || !plan || !l || !r) return fail(HB_ERR_INVALID_ARG, "hb_job_submit: null argument");
  if (plan->max_col_l >= l->n_cols || plan->max_col_r >= r->n_cols)
    return fail(HB_ERR_PLAN, "plan references column %d/%d but tables have %d/%d columns",
                plan->max_col_l, plan->max_col_r, l->n_cols, r->n_cols);
  if (l->ctx != ctx || r->ctx != ctx)
    return fail(HB_ERR_INVALID_ARG, "hb_job_submit: table from another context");
  HB_TRY(set_device(ctx));
  auto job = std::make_unique<Job>();
  Job *j = job.get();
  j->id = job_id;
  j->ctx = ctx;
  j->plan = plan;
  j->L = const_cast<hb_table *>(l);
  j->R = const_cast<hb_table *>(r);
  j->stream = ctx->streams[ctx->next_stream++ % ctx->streams.size()];
  const unsigned long long base =
      std::max<unsigned long long>(1 << 16, (unsigned long long)(l->n_rows + r->n_rows) * 2);
  j->list_cap = base * 4;
  j->out_cap = base;
  j->audit_cap = 1 << 14;
  unsigned long long sc = 1024;
  while (sc < 2 * j->out_cap) sc <<= 1;
  j->set_cap = sc;
  j->n_counters = kCntStage0;
  for (const PlanRule &rule : plan->rules) j->n_counters += 1 + 2 * (int)rule.sims.size();
  {
    void *hc = nullptr;
    HB_TRY(ctx->hpool.alloc((size_t)j->n_counters * sizeof(unsigned long long), &hc, nullptr));
    j->h_cnt = (unsigned long long *)hc;
  }
  HB_TRY(ctx->events.get(&j->ev_start));
  HB_TRY(ctx->events.get(&j->ev_end));
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->jobs.count(job_id))
      return fail(HB_ERR_INVALID_ARG, "job id %llu already in use", (unsigned long long)job_id);
    j->shape = ctx->shape;
    ctx->jobs.emplace(job_id, std::move(job));
  }
  std::lock_guard<std::mutex> lk(j->mu);
  hb_status st = enqueue_job(j);
  if (st != HB_OK) {
    j->state = JobState::kFailed;
    j->error = st;
    j->error_text = g_last_error;
  }
  return st;
}

hb_status hb_job_query(hb_ctx *ctx, uint64_t job_id) {
  Job *j = ctx ? find_job(ctx, job_id) : nullptr;
  if (!j) return fail(HB_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
  return job_poll(j, false);
}

hb_status hb_job_wait(hb_ctx *ctx, uint64_t job_id) {
  Job *j = ctx ? find_job(ctx, job_id) : nullptr;
  if (!j) return fail(HB_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
  return job_poll(j, true);
}

hb_status hb_job_results(hb_ctx *ctx, uint64_t job_id, const uint32_t **pairs, uint64_t *n_pairs,
                         const hb_audit_entry **audit, uint64_t *n_audit) {
  Job *j = ctx ? find_job(ctx, job_id) : nullptr;
  if (!j) return fail(HB_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
  HB_TRY(job_poll(j, true));
  std::lock_guard<std::mutex> lk(j->mu);
  // The kernels wrote the outputs into pinned memory; the job's completion
  // event orders those writes before this read.
  if (pairs) *pairs = j->h_out;
  if (n_pairs) *n_pairs = j->h_cnt[kCntOut];
  if (audit) *audit = j->h_audit;
  if (n_audit) *n_audit = j->h_cnt[kCntAudit];
  return HB_OK;
}

hb_status hb_job_stats_get(hb_ctx *ctx, uint64_t job_id, hb_job_stats *out) {
  Job *j = ctx ? find_job(ctx, job_id) : nullptr;
  if (!j) return fail(HB_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
  HB_TRY(job_poll(j, true));
  std::lock_guard<std::mutex> lk(j->mu);
  if (out) *out = j->stats;
  return HB_OK;
}

hb_status hb_job_release(hb_ctx *ctx, uint64_t job_id) {
  std::unique_ptr<Job> job;
  {
    if (!ctx) return fail(HB_ERR_INVALID_ARG, "hb_job_release: null context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->jobs.find(job_id);
    if (it == ctx->jobs.end())
      return fail(HB_ERR_NOT_FOUND, "unknown job %llu", (unsigned long long)job_id);
    job = std::move(it->second);
    ctx->jobs.erase(it);
  }
  HB_TRY(set_device(ctx));
  HIP_OK(hipStreamSynchronize(job->stream));
  job_free_all(job.get());
  return HB_OK;
}

hb_status hb_block(hb_ctx *ctx, const hb_plan *plan, const hb_table *l, const hb_table *r,
                   uint32_t **pairs, uint64_t *n_pairs, hb_audit_entry **audit,
                   uint64_t *n_audit, hb_job_stats *stats) {
  static std::atomic<uint64_t> seq{0};
  const uint64_t id = 0x8000000000000000ull | seq++;
  HB_TRY(hb_job_submit(ctx, plan, l, r, id));
  const uint32_t *p = nullptr;
  const hb_audit_entry *a = nullptr;
  uint64_t n = 0, na = 0;
  hb_status st = hb_job_results(ctx, id, &p, &n, &a, &na);
  if (st == HB_OK && stats) st = hb_job_stats_get(ctx, id, stats);
  if (st == HB_OK) {
    if (pairs) {
      *pairs = (uint32_t *)std::malloc(std::max<size_t>(n * 8, 8));
      if (!*pairs) st = fail(HB_ERR_OOM, "malloc");
      else if (n) std::memcpy(*pairs, p, n * 8);
    }
    if (n_pairs) *n_pairs = n;
    if (audit && st == HB_OK) {
      *audit = (hb_audit_entry *)std::malloc(std::max<size_t>(na * sizeof(hb_audit_entry), 8));
      if (!*audit) st = fail(HB_ERR_OOM, "malloc");
      else if (na) std::memcpy(*audit, a, na * sizeof(hb_audit_entry));
    }
    if (n_audit) *n_audit = na;
  }
  hb_status st2 = hb_job_release(ctx, id);
  return st != HB_OK ? st : st2;
}

void hb_free(void *p) { std::free(p); }

hb_status hb_jw_from_counts(hb_ctx *ctx, const int32_t *mstp, int64_t n, double *out,
                            double *out_fused) {
  if (!ctx || n < 0 || (n && (!mstp || !out || !out_fused)))
    return fail(HB_ERR_INVALID_ARG, "hb_jw_from_counts: bad argument");
  if (n == 0) return HB_OK;
  HB_TRY(set_device(ctx));
  for (int64_t i = 0; i < n; i++) {
    const int32_t m = mstp[4 * i], S = mstp[4 * i + 1], T = mstp[4 * i + 2], p = mstp[4 * i + 3];
    if (S < 0 || T < S || m < 0 || m > S || p < 0 || p > S)
      return fail(HB_ERR_INVALID_ARG, "hb_jw_from_counts: tuple %lld is not 0<=m,p<=S<=T",
                  (long long)i);
  }
  hipStream_t st = ctx->streams[0];
  int32_t *d_in = nullptr;
  double *d_out = nullptr;
  HIP_OK(hipMalloc(&d_in, (size_t)n * 16));
  HIP_OK(hipMalloc(&d_out, (size_t)n * 16));
  HIP_OK(hipMemcpyAsync(d_in, mstp, (size_t)n * 16, hipMemcpyHostToDevice, st));
  hb::launch_jw_from_counts(d_in, n, d_out, d_out + n, st);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, d_out, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(out_fused, d_out + n, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
  HIP_OK(hipFree(d_in));
  HIP_OK(hipFree(d_out));
  return HB_OK;
}

hb_status hb_filter_coefs(float threshold, float coef[4]) {
  if (!coef) return fail(HB_ERR_INVALID_ARG, "hb_filter_coefs: null output");
  for (int p = 0; p < 4; p++) coef[p] = filter_coef((double)threshold, p);
  return HB_OK;
}

}  // extern "C"
