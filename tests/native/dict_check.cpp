// CPU check of the runtime's dictionary coding (code_strings): equal codes
// exactly for equal byte strings, the same codes when a column is coded
// again, codes below code_bound(), on the worker pool.  Built by
// tests/test_dictionary.py from the runtime source; runs no GPU code.
#include <chrono>
#include <cstdio>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "../../hyperblocker_amd/csrc/hb_runtime.cpp"

int main() {
  hbrt::Dictionary d;
  std::mt19937_64 rng(7);
  // (rows, distinct values): high cardinality, few values, one value, empty strings
  const size_t shapes[][2] = {{100000, 400000}, {100000, 21}, {5000, 1}, {3000, 50}, {1, 1}};
  for (const auto &sh : shapes) {
    const size_t n = sh[0];
    std::vector<std::string> vals(n);
    for (size_t i = 0; i < n; i++) {
      const uint64_t v = rng() % sh[1];
      vals[i] = std::to_string(v) + (v % 7 == 0 ? std::string() : std::string("x"));
      if (v % 13 == 0) vals[i].clear();
      if (v % 17 == 0) vals[i].push_back('\x80');  // a byte above 127
    }
    std::vector<uint32_t> a(n), b(n);
    auto get = [&](size_t r) { return std::string_view(vals[r]); };
    hbrt::code_strings(d, n, get, a.data());
    hbrt::code_strings(d, n, get, b.data());
    std::map<std::string, uint32_t> by_value;
    std::map<uint32_t, std::string> by_code;
    const uint64_t bound = d.code_bound();
    for (size_t i = 0; i < n; i++) {
      if (a[i] != b[i]) return std::printf("FAIL recoding changed row %zu\n", i), 1;
      if (a[i] >= bound) return std::printf("FAIL code %u >= bound %llu\n", a[i], (unsigned long long)bound), 1;
      auto it = by_value.emplace(vals[i], a[i]).first;
      if (it->second != a[i]) return std::printf("FAIL equal strings, codes %u %u\n", it->second, a[i]), 1;
      auto jt = by_code.emplace(a[i], vals[i]).first;
      if (jt->second != vals[i]) return std::printf("FAIL code %u for two strings\n", a[i]), 1;
    }
    std::printf("rows %zu distinct %zu ok\n", n, by_value.size());
  }
  std::printf("ok\n");
  return 0;
}
