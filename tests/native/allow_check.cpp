// Exhaustive CPU check of the bound filter's allowance arithmetic
// (hyperblocker_amd/csrc/hb_allow.h, compiled here from the kernels' own
// header).  For every pair of cell lengths la, lb in 1..255, every bound
// coefficient k given on the command line (2 * coef of hb_filter_coefs over
// a grid of thresholds) and a reciprocal of s = 1/la + 1/lb perturbed by up
// to 2^-22 relative (v_rcp_f32 is within 1 ulp), the allowance decision
// "SAD > allow" equals the direct float test (float)t * s < k for every
// t = la + lb - SAD in [0, la + lb].  The direct test is monotone in t, so
// it rejects exactly the t below T* = min { t : !((float)t * s < k) }; the
// allowance rejects exactly the t below neg_allowance + tmax.  The check
// compares the two thresholds, then re-tests every t for a sample of cases.
//
// Usage: allow_check <k0> [<k1> ...]; prints "cases N mismatches M" and
// exits non-zero on a mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#define HB_ALLOW_FN static inline
#include "../../hyperblocker_amd/csrc/hb_allow.h"

namespace {

// T* of the direct test over t in [0, tmax + 1] (tmax + 1: every t rejected).
int32_t first_kept(int32_t tmax, float s, float k) {
  int32_t lo = 0, hi = tmax + 1;  // answer in [lo, hi]
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo) / 2;
    if (!((float)mid * s < k)) hi = mid; else lo = mid + 1;
  }
  return lo;
}

bool rejected_by_allowance(uint32_t nal, int32_t t, int32_t tmax) {
  const uint32_t sad = (uint32_t)(tmax - t);
  return (int32_t)(nal + sad) > 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<float> ks;
  for (int i = 1; i < argc; i++) ks.push_back(std::strtof(argv[i], nullptr));
  if (ks.empty()) {
    std::fprintf(stderr, "usage: allow_check k...\n");
    return 2;
  }
  const double rel[] = {-0x1p-22, -0x1p-23, 0.0, 0x1p-23, 0x1p-22};
  const int n_threads = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  std::vector<unsigned long long> cases(n_threads, 0), bad(n_threads, 0);
  std::vector<std::thread> th;
  for (int w = 0; w < n_threads; w++) {
    th.emplace_back([&, w] {
      for (uint32_t la = 1 + w; la <= 255; la += n_threads) {
        const float rla = 1.0f / (float)la;
        for (uint32_t lb = 1; lb <= 255; lb++) {
          const float rlb = 1.0f / (float)lb;
          const float s = rla + rlb;
          const int32_t tmax = (int32_t)(la + lb);
          for (size_t q = 0; q < ks.size(); q++) {
            const float k = ks[q];
            const int32_t want = k > 0.0f ? first_kept(tmax, s, k) : 0;
            for (double r : rel) {
              const float inv = (float)((1.0 / (double)s) * (1.0 + r));
              uint32_t nal0 = 0, nal3 = 0;
              hb::allowances_inv(la, rla, lb, rlb, k, k, inv, nal0, nal3);
              cases[w]++;
              // allowance threshold: rejected iff t < tmax + (int32)nal
              const int32_t got = k > 0.0f ? tmax + (int32_t)nal0 : 0;
              bool ok = nal0 == nal3 && (k > 0.0f ? got == want : (int32_t)nal0 < -tmax);
              // every t of a sample of cases, through the accumulator itself
              if (ok && ((la * 7 + lb * 13 + q) % 61 == 0))
                for (int32_t t = 0; t <= tmax; t++)
                  if (rejected_by_allowance(nal0, t, tmax) != ((float)t * s < k)) ok = false;
              if (!ok) {
                if (bad[w] < 5)
                  std::printf("MISMATCH la %u lb %u k %.9g rel %g: allowance T* %d direct %d\n",
                              la, lb, (double)k, r, got, want);
                bad[w]++;
              }
            }
          }
        }
      }
    });
  }
  for (auto &t : th) t.join();
  unsigned long long nc = 0, nb = 0;
  for (int w = 0; w < n_threads; w++) {
    nc += cases[w];
    nb += bad[w];
  }
  // the never-rejected and reject-all seeds
  {
    uint32_t n0 = 0, n3 = 0;
    hb::allowances_inv(0, INFINITY, 40, 1.0f / 40, 1.5f, 1.5f, 0.0f, n0, n3);
    for (int32_t t = 0; t <= 600; t++)
      if (rejected_by_allowance(n0, t, 40)) nb++;
    if ((int32_t)(hb::kRejectAll + 0u) <= 0) nb++;
  }
  std::printf("cases %llu mismatches %llu\n", nc, nb);
  return nb ? 1 : 0;
}
