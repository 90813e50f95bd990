// tamc_math.hpp -- arithmetic shared by the gfx950 kernels and the host runtime.
//
//  * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the counter-based generator that gives
//    every (sweep, group of four mutually uncoupled sites, replica) its own four 32-bit words.
//  * tamc_expf_cr: exp of a float evaluated in double precision (Cody-Waite reduction by ln 2,
//    degree-13 Taylor polynomial in Horner form with explicit fma, exact scaling by 2^k) and rounded
//    once to float.  The reference's acceptance test uses f32::exp, i.e. glibc expf; the runtime
//    pairs this function with a table of the inputs where glibc expf differs from it (generated by
//    expf_table_gen.cpp from glibc itself), so accept decisions are identical to the reference's.
//  * rand-0.8-compatible conversions: u = (w >> 8) * 2^-24.
//
// Every function is written so that host (x86-64 clang) and device (gfx950) evaluate it with the
// same IEEE operations: contraction is disabled and fused operations are spelled out.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define TAMC_HD __host__ __device__ __forceinline__
#else
#define TAMC_HD inline
#endif

#include <math.h>

namespace tamc {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

// Stream tags carried in the fourth counter word.
enum : uint32_t { kTagSweep = 0, kTagExchange = 1, kTagIcm = 2, kTagInit = 3, kTagPercolation = 4 };

struct U4 {
  uint32_t x, y, z, w;
};

// a ^ b ^ c; on gfx950 one v_bitop3_b32 (truth table 0x96) instead of two v_xor_b32
TAMC_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

TAMC_HD U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
    const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
    const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return U4{c0, c1, c2, c3};
}

// word (lane & 3) of v, computed with masks so that a run-time lane never becomes a memory lookup
TAMC_HD uint32_t pick(const U4 &v, uint32_t lane) {
  const uint32_t l = lane & 3u;
  return (v.x & (0u - (uint32_t)(l == 0u))) | (v.y & (0u - (uint32_t)(l == 1u))) |
         (v.z & (0u - (uint32_t)(l == 2u))) | (v.w & (0u - (uint32_t)(l == 3u)));
}

// rand 0.8 Standard f32 from one 32-bit word, as an integer numerator over 2^24.
TAMC_HD uint32_t u24_bits(uint32_t w) { return w >> 8; }
TAMC_HD float u24(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }

// exp(x) for float x, evaluated in double and rounded once.  Valid for x <= 0.
TAMC_HD float expf_cr(float xf) {
#pragma clang fp contract(off)
  if (!(xf >= -104.0f)) return (xf != xf) ? xf : 0.0f;  // NaN propagates; far underflow is +0
  if (xf == 0.0f) return 1.0f;
  const double x = (double)xf;
  const double kInvLn2 = 1.4426950408889634;
  const double kLn2Hi = 0.6931471806019545;      // ln 2 to 32 significant bits: k*kLn2Hi is exact
  const double kLn2Lo = -4.2009150726810846e-11;  // ln 2 - kLn2Hi
  const double kd = rint(x * kInvLn2);
  const double r = fma(-kd, kLn2Lo, fma(-kd, kLn2Hi, x));
  double p = 1.6059043836821613e-10;  // 1/13!
  p = fma(p, r, 2.08767569878681e-09);
  p = fma(p, r, 2.505210838544172e-08);
  p = fma(p, r, 2.755731922398589e-07);
  p = fma(p, r, 2.7557319223985893e-06);
  p = fma(p, r, 2.48015873015873e-05);
  p = fma(p, r, 0.0001984126984126984);
  p = fma(p, r, 0.001388888888888889);
  p = fma(p, r, 0.008333333333333333);
  p = fma(p, r, 0.041666666666666664);
  p = fma(p, r, 0.16666666666666666);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  // scale by 2^k in two exact steps (k in [-151, 0])
  const int k = (int)kd;
  const int k1 = k / 2, k2 = k - k1;
  union {
    uint64_t u;
    double d;
  } s1, s2;
  s1.u = (uint64_t)(1023 + k1) << 52;
  s2.u = (uint64_t)(1023 + k2) << 52;
  return (float)((p * s1.d) * s2.d);
}

// Table of inputs where glibc expf differs from expf_cr: sorted float bit patterns and the glibc
// result bits.  Looked up only when a uniform falls within one ulp of the computed probability.
struct ExpfExceptions {
  const uint32_t *x_bits;
  const uint32_t *y_bits;
  uint32_t count;
};

TAMC_HD uint32_t f2u(float f) {
  union {
    float f;
    uint32_t u;
  } c;
  c.f = f;
  return c.u;
}
TAMC_HD float u2f(uint32_t u) {
  union {
    float f;
    uint32_t u;
  } c;
  c.u = u;
  return c.f;
}

TAMC_HD float expf_ref(float x, const ExpfExceptions &tab) {
  float y = expf_cr(x);
  const uint32_t xb = f2u(x);
  uint32_t lo = 0, hi = tab.count;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    const uint32_t v = tab.x_bits[mid];
    if (v == xb) return u2f(tab.y_bits[mid]);
    if (v < xb) lo = mid + 1;
    else hi = mid;
  }
  return y;
}

// Exact "u < expf(x)" with u = k * 2^-24, using the table only when u is within one ulp of the
// correctly rounded probability.
TAMC_HD bool accept_exact(float x, uint32_t k, const ExpfExceptions &tab) {
  const float u = (float)k * (1.0f / 16777216.0f);
  const float p = expf_cr(x);
  const float up = u2f(f2u(p) + 1u), dn = (p > 0.0f) ? u2f(f2u(p) - 1u) : 0.0f;
  if (u < dn) return true;
  if (u > up) return false;
  return u < expf_ref(x, tab);
}

}  // namespace tamc
