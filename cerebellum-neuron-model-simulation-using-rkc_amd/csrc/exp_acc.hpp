// exp(x) rounded to nearest from a double-double evaluation, for the model
// code whose trajectories amplify last-bit differences (the gap-coupled
// inferior-olive grid).
//
// Method (Cody & Waite argument reduction, then Taylor series with repeated
// squaring, all in double-double arithmetic with FMA error-free products,
// Dekker/Knuth two-sum):
//   k = rint(x / ln 2),  r = x - k ln 2           (ln 2 split into 3 parts)
//   exp(r) = (exp(r / 2^8))^(2^8),  exp(s) = sum_{i <= 10} s^i / i!
//   exp(x) = 2^k exp(r)
// The double-double result carries ~100 bits, so rounding its sum to double
// gives the correctly rounded exp except in cases closer to a rounding
// boundary than ~2^-45 ulp.  Arguments outside (-708, 709) fall back to
// exp(), which is exact at the overflow / underflow ends that matter here.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace cbrkc {

struct dd {
  double hi, lo;
};

__host__ __device__ inline dd dd_two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ inline dd dd_fast_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
__host__ __device__ inline dd dd_add(dd a, dd b) {
  dd s = dd_two_sum(a.hi, b.hi);
  dd t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}
__host__ __device__ inline dd dd_mul(dd a, dd b) {
  const double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e = fma(a.hi, b.lo, e);
  e = fma(a.lo, b.hi, e);
  return dd_fast_two_sum(p, e);
}
__host__ __device__ inline dd dd_mul_d(dd a, double b) {
  const double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e = fma(a.lo, b, e);
  return dd_fast_two_sum(p, e);
}

__host__ __device__ inline double exp_acc(double x) {
  if (!(x > -708.0 && x < 709.0)) return exp(x);
  // ln 2 = L1 + L2 + L3: L1 is ln 2 rounded to 32 fraction bits, so k L1 is
  // exact for |k| < 2^21 and x - k L1 is exact (Sterbenz); L2, L3 are the next
  // two parts rounded to double
  const double L1 = 0.6931471806019545;
  const double L2 = -4.2009150726810846e-11;
  const double L3 = -1.3124698417785255e-27;
  const double k = rint(x * 1.4426950408889634074);
  const double t = x - k * L1;                 // exact
  // r = t - k L2 - k L3 in double-double
  const double p = k * L2;
  const double pe = fma(k, L2, -p);
  dd r = dd_two_sum(t, -p);
  r.lo -= pe + k * L3;
  r = dd_fast_two_sum(r.hi, r.lo);
  // s = r / 2^8 (exact scaling)
  const dd s = {r.hi * 0.00390625, r.lo * 0.00390625};
  // Taylor sum for exp(s) - 1 in Horner form, |s| < 0.0014: 10 terms reach 2^-110
  dd acc = {2.755731922398589e-07, 0.0};   // 1/10!
  // 1/i! for i = 9..0 as double-double: the double nearest and the exact
  // remainder 1/i! - double(1/i!) rounded to double
  const double inv[10] = {2.7557319223985893e-06, 2.48015873015873e-05, 0.0001984126984126984,
                          0.001388888888888889,   0.008333333333333333, 0.041666666666666664,
                          0.16666666666666666,    0.5,                  1.0,
                          1.0};
  const double inv_lo[10] = {-1.858393274046472e-22, 2.1511947866775882e-23, 1.7209558293420705e-22,
                             -5.300543954373577e-20, 1.1564823173178714e-19, 2.3129646346357427e-18,
                             9.25185853854297e-18,   0.0,                    0.0,
                             0.0};
  for (int i = 0; i < 10; i++) {
    acc = dd_mul(acc, s);
    acc = dd_add(acc, dd{inv[i], inv_lo[i]});
  }
  // square 8 times: exp(r) = exp(s)^256
  for (int i = 0; i < 8; i++) acc = dd_mul(acc, acc);
  return ldexp(acc.hi + acc.lo, (int)k);
}

// The models' exponential (inferior olive, Golgi cells): the host C
// library's on the host (initial state), the correctly rounded exp_acc on the
// device.  These populations amplify last-bit differences of exp into
// spike-time shifts over hundreds of milliseconds, so the device follows a
// correctly rounded exp rather than the device library's (faithful, not
// correctly rounded) one.
// CB_HOST_EMU_DEVICE / CB_HOST_EMU_HOST: the host build of the device code
// used by the tests (a host stand-in for the HIP header, one thread per
// lane) takes the device's or the host's exponential explicitly.
#if defined(CB_HOST_EMU_DEVICE)
static inline double mexp(double x) { return exp_acc(x); }
#elif defined(CB_HOST_EMU_HOST)
static inline double mexp(double x) { return ::exp(x); }
#else
__host__ inline double mexp(double x) { return ::exp(x); }
__device__ inline double mexp(double x) { return exp_acc(x); }
#endif

}  // namespace cbrkc
