// Seeded synthetic entity-resolution tables for the blocking workloads.
//
// A table pair (L, R) shares one schema.  The first n_dup rows of R are
// perturbed copies of distinct L rows: in every column marked "perturb",
// each character independently, with probability edit_rate, is substituted,
// followed by an inserted character, deleted, or swapped with its successor.
// All other R rows are fresh draws.  Every random choice comes from a
// splitmix64 stream keyed by (seed, side, row), so the output depends only
// on the arguments and is identical on every x86-64 host.
//
// Characters of free text are printable ASCII (0x20-0x7E) without ',' and '"',
// so CSV files written from these tables need no quoting.
#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  int64_t range(int64_t lo, int64_t hi) {  // inclusive
    return lo + (int64_t)(next() % (uint64_t)(hi - lo + 1));
  }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

uint64_t mix(uint64_t a, uint64_t b) {
  Rng r(a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull));
  r.next();
  return r.next();
}

const char kAlphabet[] =
    " !#$%&'()*+-./0123456789:;<=>?@ABCDEFGHIJKLMNOPQRSTUVWXYZ[\\]^_`"
    "abcdefghijklmnopqrstuvwxyz{|}~";
const int kAlphabetSize = (int)sizeof(kAlphabet) - 1;  // 93

const char *kVenues[20] = {
    "SIGMOD Conference", "VLDB", "ICDE", "SIGMOD Record", "VLDB J.",
    "TODS", "KDD", "CIKM", "EDBT", "PODS", "WWW", "SIGIR", "ICDT",
    "DASFAA", "SSDBM", "TKDE", "WSDM", "ER", "DEXA", "Inf. Syst."};

}  // namespace

extern "C" {

// Column kinds.
enum { HBG_ID = 0, HBG_TEXT = 1, HBG_VENUE = 2, HBG_INT = 3, HBG_DIGITS = 4 };

typedef struct {
  int32_t kind;
  int32_t min_len, max_len;  // TEXT: length bounds; DIGITS: digit count
  double log_mu, log_sigma;  // TEXT: lognormal length if log_sigma > 0
  int32_t lo, hi;            // INT: inclusive value range
  int32_t perturb;           // R copies get character edits in this column
} hbg_col;

// Width (bytes per cell, NUL included) that hbg_generate writes for a column.
int64_t hbg_width(const hbg_col *c) {
  int64_t w;
  switch (c->kind) {
  case HBG_ID: w = 21; break;
  case HBG_TEXT: w = c->max_len + 1 + (c->perturb ? 16 : 0); break;
  case HBG_VENUE: w = 18; break;
  case HBG_INT: w = 12; break;
  case HBG_DIGITS: w = c->max_len + 1; break;
  default: w = 1;
  }
  return (w + 3) & ~int64_t(3);
}

static void gen_cell(const hbg_col &c, Rng &rng, int64_t row, std::string &out) {
  out.clear();
  switch (c.kind) {
  case HBG_ID: out = std::to_string(row); break;
  case HBG_TEXT: {
    int64_t len;
    if (c.log_sigma > 0) {
      double v = std::exp(c.log_mu + c.log_sigma * rng.normal());
      len = (int64_t)std::llround(v);
      if (len < c.min_len) len = c.min_len;
      if (len > c.max_len) len = c.max_len;
    } else {
      len = rng.range(c.min_len, c.max_len);
    }
    for (int64_t i = 0; i < len; i++) out.push_back(kAlphabet[rng.next() % kAlphabetSize]);
    break;
  }
  case HBG_VENUE: out = kVenues[rng.next() % 20]; break;
  case HBG_INT: out = std::to_string(rng.range(c.lo, c.hi)); break;
  case HBG_DIGITS:
    for (int i = 0; i < c.max_len; i++) out.push_back((char)('0' + rng.next() % 10));
    break;
  }
}

static void perturb(const std::string &in, double rate, Rng &rng, size_t cap, std::string &out) {
  out.clear();
  for (size_t i = 0; i < in.size(); i++) {
    if (rng.uniform() < rate) {
      switch (rng.next() % 4) {
      case 0: out.push_back(kAlphabet[rng.next() % kAlphabetSize]); break;  // substitute
      case 1:  // insert after
        out.push_back(in[i]);
        out.push_back(kAlphabet[rng.next() % kAlphabetSize]);
        break;
      case 2: break;  // delete
      case 3:         // swap with the next character
        if (i + 1 < in.size()) {
          out.push_back(in[i + 1]);
          out.push_back(in[i]);
          i++;
        } else {
          out.push_back(in[i]);
        }
        break;
      }
    } else {
      out.push_back(in[i]);
    }
  }
  if (out.size() > cap) out.resize(cap);
}

// Fills caller-allocated fixed-width column buffers (n_rows * hbg_width each,
// zero-initialised by the caller).  dup_src receives, for each of the first
// n_dup R rows, the L row it copies.  Returns 0, or -1 on bad arguments.
int hbg_generate(uint64_t seed, int32_t n_cols, const hbg_col *cols, int64_t n_l,
                 int64_t n_r, int64_t n_dup, double edit_rate, char **out_l,
                 char **out_r, int64_t *dup_src, int32_t n_threads) {
  if (n_dup > n_l || n_dup > n_r || n_cols <= 0) return -1;
  std::vector<int64_t> widths(n_cols);
  for (int c = 0; c < n_cols; c++) widths[c] = hbg_width(&cols[c]);
  // Distinct source rows for the copies: a seeded partial Fisher-Yates.
  {
    std::vector<int64_t> perm(n_l);
    for (int64_t i = 0; i < n_l; i++) perm[i] = i;
    Rng rng(mix(seed, 0xD0D0));
    for (int64_t i = 0; i < n_dup; i++) {
      int64_t j = rng.range(i, n_l - 1);
      std::swap(perm[i], perm[j]);
      dup_src[i] = perm[i];
    }
  }
  if (n_threads < 1) n_threads = 1;
  auto fill_l = [&](int64_t r0, int64_t r1) {
    std::string cellv;
    for (int64_t r = r0; r < r1; r++) {
      Rng rng(mix(seed * 2 + 0, (uint64_t)r));
      for (int c = 0; c < n_cols; c++) {
        gen_cell(cols[c], rng, r, cellv);
        std::memcpy(out_l[c] + r * widths[c], cellv.data(), cellv.size());
      }
    }
  };
  auto fill_r = [&](int64_t r0, int64_t r1) {
    std::string cellv, src, pert;
    for (int64_t r = r0; r < r1; r++) {
      Rng rng(mix(seed * 2 + 1, (uint64_t)r));
      for (int c = 0; c < n_cols; c++) {
        if (r < n_dup && cols[c].kind != HBG_ID) {
          const char *s = out_l[c] + dup_src[r] * widths[c];
          src.assign(s, strnlen(s, (size_t)widths[c]));
          if (cols[c].perturb) {
            perturb(src, edit_rate, rng, (size_t)widths[c] - 1, pert);
            std::memcpy(out_r[c] + r * widths[c], pert.data(), pert.size());
          } else {
            std::memcpy(out_r[c] + r * widths[c], src.data(), src.size());
          }
        } else {
          gen_cell(cols[c], rng, r, cellv);
          std::memcpy(out_r[c] + r * widths[c], cellv.data(), cellv.size());
        }
      }
    }
  };
  auto run = [&](int64_t n, const std::function<void(int64_t, int64_t)> &f) {
    std::vector<std::thread> th;
    int64_t per = (n + n_threads - 1) / n_threads;
    for (int t = 0; t < n_threads; t++) {
      int64_t a = t * per, b = std::min<int64_t>(n, a + per);
      if (a < b) th.emplace_back(f, a, b);
    }
    for (auto &x : th) x.join();
  };
  run(n_l, fill_l);  // R copies read L, so L is complete first
  run(n_r, fill_r);
  return 0;
}

// Writes a table as CSV without a header row (the reference reads every
// line as data), one line per row, cells joined by sep.
int hbg_write_csv(const char *path, int32_t n_cols, int64_t n_rows, char *const *cols,
                  const int64_t *widths, char sep) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return -1;
  std::vector<char> buf;
  buf.reserve(1 << 20);
  for (int64_t r = 0; r < n_rows; r++) {
    for (int c = 0; c < n_cols; c++) {
      const char *s = cols[c] + r * widths[c];
      size_t n = strnlen(s, (size_t)widths[c]);
      buf.insert(buf.end(), s, s + n);
      buf.push_back(c + 1 < n_cols ? sep : '\n');
    }
    if (buf.size() > (1 << 20)) {
      std::fwrite(buf.data(), 1, buf.size(), f);
      buf.clear();
    }
  }
  std::fwrite(buf.data(), 1, buf.size(), f);
  return std::fclose(f) == 0 ? 0 : -1;
}

}  // extern "C"
