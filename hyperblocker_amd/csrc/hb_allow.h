// Bound-filter allowances: the float test of the filter in integer form.
//
// The filter rejects a candidate pair of cell lengths la, lb when
//   (float)t * s < k,   s = rla + rlb (float reciprocal lengths),
//                       t = la + lb - SAD = 2 * (sum over byte classes of
//                           the smaller count), k = 2 * coef
// (the Jaro bound of device_func.cuh:127-128 with the Winkler bonus of
// :229 bounded by the prefix class).  For fixed lengths the left side does
// not decrease with t, so the rejected t are exactly those below
//   T* = min { t : !((float)t * s < k) },
// i.e. the pair is rejected iff SAD > allow = la + lb - T*.  neg_allowance
// returns -allow (two's complement) found by that same float comparison;
// seeding the SAD accumulator with it leaves SAD - allow, positive iff the
// pair is rejected.
//
// Shared by the kernels (hb_kernels.hip) and the exhaustive CPU check of
// tests/native/allow_check.cpp, so that the check runs this code.  The
// includer defines HB_ALLOW_FN (the function qualifiers) first; the
// reciprocal of s is an argument, formed by the caller (v_rcp_f32 on the
// device: within 1 ulp of 1/s).
#ifndef HB_ALLOW_H_
#define HB_ALLOW_H_

#include <stdint.h>

#ifndef HB_ALLOW_FN
#error "define HB_ALLOW_FN before including hb_allow.h"
#endif

namespace hb {

// Negated allowance seed of a pair that is never rejected (the
// accumulator stays negative) and of a padding row (it stays positive).
constexpr int32_t kPassAllow = 1 << 30;
constexpr uint32_t kRejectAll = 1u << 30;

// -allow for lengths summing to tmax, s = rla + rlb, k > 0, q ~ k / s.
// t * s crosses k within 1 of ceil(q): |k / s| <= 511 and q and every
// float operation are off by at most 2^-22 relative, so one step either
// way lands on T* (clamped to [1, tmax + 1]; T* >= 1 as k > 0).
HB_ALLOW_FN uint32_t neg_allowance(int32_t tmax, float s, float q, float k) {
  if (!(k > 0.0f)) return (uint32_t)(-kPassAllow);
  int32_t t = (int32_t)__builtin_ceilf(q);
  t = t < 1 ? 1 : t;
  t = t > tmax + 1 ? tmax + 1 : t;
  if (!((float)(t - 1) * s < k))
    t -= 1;
  else if (t <= tmax && (float)t * s < k)
    t += 1;
  return (uint32_t)(t - tmax);
}

// Negated allowances of one A row for one B length, for the two prefix
// classes (k0: first bytes differ; k3: a later one of the first 4 bytes
// differs).  inv approximates 1 / (rla + rlb).  An empty or over-long cell
// (reciprocal +inf) is never rejected.
HB_ALLOW_FN void allowances_inv(uint32_t la, float rla, uint32_t lb, float rlb, float k0,
                                float k3, float inv, uint32_t &nal0, uint32_t &nal3) {
  const float s = rla + rlb;
  if (!(s < __builtin_inff())) {
    nal0 = nal3 = (uint32_t)(-kPassAllow);
    return;
  }
  const int32_t tmax = (int32_t)((la & 0xffffu) + (lb & 0xffffu));
  nal0 = neg_allowance(tmax, s, k0 * inv, k0);
  nal3 = neg_allowance(tmax, s, k3 * inv, k3);
}

}  // namespace hb

#endif  // HB_ALLOW_H_
