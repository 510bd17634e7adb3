// run_hyperblocker_exec -- drop-in for examples/run_hyperblocker.cu:22-59.
//
// Same flags (gflags syntax: -flag=value, --flag=value, -flag value, bool
// -flag / -noflag): -data_l -data_r -rule_dir -o -sep -scheduler
// -read_header -n_partitions -prefix_hash_pred_idx.  -read_header is parsed
// and ignored and -scheduler only selects among equivalent single-GPU
// placements, as in the reference (core/hyperblocker.cuh:85-86; one device).
//
// Additional flags: -append (open the output in append mode like
// host_reducer.h:83; default truncates), -audit <path>, -device <n>,
// -metrics <path> (JSON), -max_inflight <n> (partitions in flight).
//
// Flow = HyperBlocker::Run (core/hyperblocker.cuh:103-131): plan from the
// rule directory, both CSVs, hash partitioning, then per partition an upload
// and an asynchronous job; completed jobs are written in submission order.
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hyperblocker_amd.h"

namespace {

struct Flag {
  enum Type { kString, kBool, kUint64, kInt } type;
  std::string sval;
  bool bval = false;
  unsigned long long uval = 0;
  long long ival = 0;
  const char *help;
};

std::map<std::string, Flag> g_flags = {
    {"data_l", {Flag::kString, "", false, 0, 0, "input dir 2."}},
    {"data_r", {Flag::kString, "", false, 0, 0, "input dir 1."}},
    {"rule_dir", {Flag::kString, "", false, 0, 0, "root path of rules."}},
    {"o", {Flag::kString, "", false, 0, 0, "output path."}},
    {"sep", {Flag::kString, ",", false, 0, 0, "separator to split a line of csv file."}},
    {"scheduler", {Flag::kString, "CHBL", false, 0, 0, "scheduler type."}},
    {"read_header", {Flag::kBool, "", false, 0, 0, "whether to read header of csv."}},
    {"n_partitions", {Flag::kUint64, "", false, 1, 0, "number of partitions."}},
    {"prefix_hash_pred_idx", {Flag::kUint64, "", false, INT_MAX, 0, "number of partitions."}},
    {"append", {Flag::kBool, "", false, 0, 0, "append to -o instead of truncating it."}},
    {"audit", {Flag::kString, "", false, 0, 0, "write near-threshold evaluations here."}},
    {"device", {Flag::kInt, "", false, 0, 0, "GPU ordinal."}},
    {"metrics", {Flag::kString, "", false, 0, 0, "write a JSON run summary here."}},
    {"max_inflight", {Flag::kInt, "", false, 0, 4, "partitions in flight on the GPU."}},
};

bool set_flag(const std::string &name, const std::string *value, bool negated) {
  auto it = g_flags.find(name);
  if (it == g_flags.end()) {
    std::fprintf(stderr, "ERROR: unknown command line flag '%s'\n", name.c_str());
    return false;
  }
  Flag &f = it->second;
  switch (f.type) {
  case Flag::kBool:
    if (!value) {
      f.bval = !negated;
    } else if (*value == "true" || *value == "1" || *value == "yes" || *value == "t" ||
               *value == "y") {
      f.bval = true;
    } else if (*value == "false" || *value == "0" || *value == "no" || *value == "f" ||
               *value == "n") {
      f.bval = false;
    } else {
      std::fprintf(stderr, "ERROR: illegal value '%s' specified for bool flag '%s'\n",
                   value->c_str(), name.c_str());
      return false;
    }
    return true;
  case Flag::kString:
    f.sval = *value;
    return true;
  case Flag::kUint64: {
    char *e = nullptr;
    errno = 0;
    unsigned long long v = std::strtoull(value->c_str(), &e, 0);
    if (value->empty() || *e || errno || (*value)[0] == '-') {
      std::fprintf(stderr, "ERROR: illegal value '%s' specified for uint64 flag '%s'\n",
                   value->c_str(), name.c_str());
      return false;
    }
    f.uval = v;
    return true;
  }
  case Flag::kInt: {
    char *e = nullptr;
    errno = 0;
    long long v = std::strtoll(value->c_str(), &e, 0);
    if (value->empty() || *e || errno) {
      std::fprintf(stderr, "ERROR: illegal value '%s' specified for int flag '%s'\n",
                   value->c_str(), name.c_str());
      return false;
    }
    f.ival = v;
    return true;
  }
  }
  return false;
}

bool parse_flags(int argc, char **argv) {
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--") break;
    if (a.size() < 2 || a[0] != '-') continue;  // positional arguments are ignored
    std::string body = a.substr(a[1] == '-' ? 2 : 1);
    if (body == "help" || body == "helpfull") {
      std::printf("run_hyperblocker_exec flags:\n");
      for (auto &kv : g_flags) std::printf("  -%s  %s\n", kv.first.c_str(), kv.second.help);
      std::exit(0);
    }
    size_t eq = body.find('=');
    if (eq != std::string::npos) {
      std::string name = body.substr(0, eq), val = body.substr(eq + 1);
      if (!set_flag(name, &val, false)) return false;
      continue;
    }
    auto it = g_flags.find(body);
    if (it == g_flags.end() && body.rfind("no", 0) == 0) {
      auto nit = g_flags.find(body.substr(2));
      if (nit != g_flags.end() && nit->second.type == Flag::kBool) {
        if (!set_flag(body.substr(2), nullptr, true)) return false;
        continue;
      }
    }
    if (it != g_flags.end() && it->second.type == Flag::kBool) {
      if (!set_flag(body, nullptr, false)) return false;
      continue;
    }
    if (it == g_flags.end()) return set_flag(body, nullptr, false);
    if (i + 1 >= argc) {
      std::fprintf(stderr, "ERROR: flag '%s' is missing its argument\n", body.c_str());
      return false;
    }
    std::string val = argv[++i];
    if (!set_flag(body, &val, false)) return false;
  }
  return true;
}

int die(const char *what) {
  std::fprintf(stderr, "run_hyperblocker_exec: %s: %s\n", what, hb_last_error());
  return 1;
}

struct InFlight {
  uint64_t id;
  int64_t part;
  hb_table *l, *r;
};

}  // namespace

int main(int argc, char **argv) {
  if (!parse_flags(argc, argv)) return 1;
  const std::string data_l = g_flags["data_l"].sval, data_r = g_flags["data_r"].sval;
  const std::string rule_dir = g_flags["rule_dir"].sval, out_path = g_flags["o"].sval;
  const std::string sep = g_flags["sep"].sval;
  const uint64_t n_partitions = g_flags["n_partitions"].uval;
  const int64_t prefix_idx = (int64_t)(int)g_flags["prefix_hash_pred_idx"].uval;
  const bool append = g_flags["append"].bval;
  const std::string audit_path = g_flags["audit"].sval, metrics_path = g_flags["metrics"].sval;
  const int device = (int)g_flags["device"].ival;
  const int max_inflight = std::max<long long>(1, g_flags["max_inflight"].ival);
  if (data_l.empty() || data_r.empty() || rule_dir.empty() || out_path.empty()) {
    std::fprintf(stderr, "run_hyperblocker_exec: -data_l, -data_r, -rule_dir and -o are required\n");
    return 1;
  }
  if (sep.empty()) {
    std::fprintf(stderr, "run_hyperblocker_exec: -sep must not be empty\n");
    return 1;
  }
  const auto t0 = std::chrono::steady_clock::now();

  // The device context (HIP runtime start-up) needs nothing from the input
  // files: it is created on a thread of its own while they load.  The
  // thread is joined before the context is used and on every early return.
  hb_ctx *ctx = nullptr;
  hb_status ctx_st = HB_OK;
  std::string ctx_err;
  std::thread ctx_init([&] {
    ctx_st = hb_ctx_create(device, &ctx);
    if (ctx_st != HB_OK) ctx_err = hb_last_error();  // the error text is per thread
  });
  struct Joiner {
    std::thread &t;
    ~Joiner() {
      if (t.joinable()) t.join();
    }
  } ctx_join{ctx_init};

  hb_host_plan *hplan = nullptr;
  if (hb_rules_load(rule_dir.c_str(), &hplan) != HB_OK) return die("rules");
  hb_serialized_plan sp;
  hb_host_plan_view(hplan, &sp);
  std::printf("SerializedExecutionPlan - n_rules: %d length: %d\n", sp.n_rules, sp.length);

  hb_host_table *L = nullptr, *R = nullptr;
  if (hb_csv_load(data_l.c_str(), sep[0], &L) != HB_OK) return die("data_l");
  if (hb_csv_load(data_r.c_str(), sep[0], &R) != HB_OK) return die("data_r");
  hb_table_view vl, vr;
  hb_host_table_view(L, &vl);
  hb_host_table_view(R, &vr);
  const auto t_load = std::chrono::steady_clock::now();

  hb_partitions *parts = nullptr;
  if (hb_partition(L, R, &sp, n_partitions, prefix_idx, &parts) != HB_OK) return die("partition");
  const int64_t n_parts = hb_partitions_count(parts);
  std::printf("Actual number of partitions: %lld\n", (long long)n_parts);
  const auto t_part = std::chrono::steady_clock::now();

  ctx_init.join();
  const auto t_ctx = std::chrono::steady_clock::now();
  if (ctx_st != HB_OK) {
    std::fprintf(stderr, "run_hyperblocker_exec: device: %s\n", ctx_err.c_str());
    return 1;
  }
  hb_plan *plan = nullptr;
  if (hb_plan_create(ctx, &sp, &plan) != HB_OK) return die("plan");

  if (!append) {
    FILE *f = std::fopen(out_path.c_str(), "wb");
    if (!f) {
      std::fprintf(stderr, "run_hyperblocker_exec: cannot open '%s'\n", out_path.c_str());
      return 1;
    }
    std::fclose(f);
  }
  FILE *audit_f = nullptr;
  if (!audit_path.empty()) {
    audit_f = std::fopen(audit_path.c_str(), "wb");
    if (!audit_f) {
      std::fprintf(stderr, "run_hyperblocker_exec: cannot open '%s'\n", audit_path.c_str());
      return 1;
    }
  }

  unsigned long long tot_pairs = 0, tot_matches = 0, tot_jw = 0, tot_cand = 0, tot_audit = 0;
  double dev_ms = 0, upload_s = 0, write_s = 0;
  auto since = [](std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
  };
  std::deque<InFlight> q;
  auto drain_one = [&]() -> int {
    InFlight f = q.front();
    q.pop_front();
    const uint32_t *pairs = nullptr;
    const hb_audit_entry *au = nullptr;
    uint64_t n = 0, na = 0;
    if (hb_job_results(ctx, f.id, &pairs, &n, &au, &na) != HB_OK) return die("job");
    hb_job_stats st;
    if (hb_job_stats_get(ctx, f.id, &st) != HB_OK) return die("job stats");
    const int64_t *lr = nullptr, *rr = nullptr;
    int64_t nl = 0, nr = 0;
    hb_partitions_get(parts, f.part, &lr, &nl, &rr, &nr);
    const auto tw = std::chrono::steady_clock::now();
    if (hb_write_pairs(out_path.c_str(), 1, L, R, lr, rr, pairs, n) != HB_OK) return die("output");
    write_s += since(tw);
    if (audit_f)
      for (uint64_t i = 0; i < na; i++)
        std::fprintf(audit_f, "%s,%s,%d,%a,%d\n", hb_host_table_cell(L, lr[au[i].l], 0),
                     hb_host_table_cell(R, rr[au[i].r], 0), au[i].slot, au[i].jw, au[i].flags);
    std::printf("Ball id %lld output: %llu\n", (long long)f.part, (unsigned long long)n);
    tot_pairs += st.pairs;
    tot_matches += n;
    tot_jw += st.jw_exact;
    tot_cand += st.candidates;
    tot_audit += na;
    dev_ms += st.device_ms;
    hb_job_release(ctx, f.id);
    hb_table_release(ctx, f.l);
    hb_table_release(ctx, f.r);
    return 0;
  };
  for (int64_t k = 0; k < n_parts; k++) {
    const int64_t *lr, *rr;
    int64_t nl, nr;
    hb_partitions_get(parts, k, &lr, &nl, &rr, &nr);
    hb_table *tl = nullptr, *tr = nullptr;
    const auto tu = std::chrono::steady_clock::now();
    if (hb_table_upload_rows(ctx, &vl, lr, nl, &tl) != HB_OK) return die("upload");
    if (hb_table_upload_rows(ctx, &vr, rr, nr, &tr) != HB_OK) return die("upload");
    upload_s += since(tu);
    if (hb_job_submit(ctx, plan, tl, tr, (uint64_t)k) != HB_OK) return die("submit");
    q.push_back({(uint64_t)k, k, tl, tr});
    if ((int)q.size() >= max_inflight)
      if (drain_one()) return 1;
  }
  while (!q.empty())
    if (drain_one()) return 1;
  if (audit_f) std::fclose(audit_f);
  const auto t_end = std::chrono::steady_clock::now();
  auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
  std::printf("HyperBlocker.Run() pairs: %llu matches: %llu load: %.3fs partition: %.3fs "
              "blocking: %.3fs total: %.3fs\n",
              tot_pairs, tot_matches, secs(t0, t_load), secs(t_load, t_part),
              secs(t_part, t_end), secs(t0, t_end));
  if (!metrics_path.empty()) {
    FILE *m = std::fopen(metrics_path.c_str(), "wb");
    if (m) {
      std::fprintf(m,
                   "{\"partitions\": %lld, \"pairs\": %llu, \"candidates\": %llu, "
                   "\"jw_exact\": %llu, \"matches\": %llu, \"audit\": %llu, "
                   "\"device_ms\": %.6f, \"load_s\": %.6f, \"partition_s\": %.6f, "
                   "\"blocking_s\": %.6f, \"total_s\": %.6f, \"ctx_wait_s\": %.6f, "
                   "\"upload_s\": %.6f, \"write_s\": %.6f}\n",
                   (long long)n_parts, tot_pairs, tot_cand, tot_jw, tot_matches, tot_audit,
                   dev_ms, secs(t0, t_load), secs(t_load, t_part), secs(t_part, t_end),
                   secs(t0, t_end), secs(t_part, t_ctx), upload_s, write_s);
      std::fclose(m);
    }
  }
  hb_plan_release(ctx, plan);
  hb_ctx_destroy(ctx);
  hb_partitions_free(parts);
  hb_host_table_free(L);
  hb_host_table_free(R);
  hb_host_plan_free(hplan);
  return 0;
}
