// Correctly rounded exp, log, pow and expm1 for the device.
//
// The reference evaluates these with the host C library (glibc 2.28+), whose
// exp/log/pow are correctly rounded in all but very rare cases.  The device
// library's versions are faithful (within 1 ulp) and differ from glibc's in a
// few percent of arguments (up to a quarter for pow), which moves positions by
// one ulp and, at knife edges (a particle exactly on a sigma plane or an
// element edge), changes a layer or element id.  These versions evaluate the
// function in double-double arithmetic (about 2^-100 relative) and round once,
// so they return the correctly rounded result and agree with glibc wherever
// glibc is correctly rounded.
//
//   exp:   x = k ln2 + j/64 + s, |s| <= 1/128; e^s by its Taylor series (13
//          terms), times the tabulated e^(j/64), scaled by 2^k
//   log:   |x - 1| < 1/32: 2 atanh(t / (2 + t)) series; otherwise one Newton
//          step y0 + log1p(x e^-y0 - 1) from the library log y0, with e^-y0
//          in double-double
//   pow:   exp of the double-double product y * log(x)
//   expm1: |x| < 1/32 Taylor series, otherwise exp(x) - 1 in double-double
// Arguments whose results are subnormal, zero, infinite or NaN, and negative
// bases of pow, go to the device library (the step never produces them).
// Double-double algorithms: Dekker's two-product (with a fused multiply-add)
// and Knuth's two-sum.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace cyl {

struct DD {
  double hi, lo;
};

}  // namespace cyl

#include "cyl_crmath_tables.h"

namespace cyl {

__device__ __forceinline__ DD dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

__device__ __forceinline__ DD dd_fast_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}

__device__ __forceinline__ DD dd_two_prod(double a, double b) {
  double p = a * b;
  return {p, fma(a, b, -p)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  DD s = dd_two_sum(a.hi, b.hi);
  DD t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ DD dd_mul(DD a, DD b) {
  DD p = dd_two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return dd_fast_two_sum(p.hi, p.lo);
}

__device__ __forceinline__ DD dd_mul_d(DD a, double b) {
  DD p = dd_two_prod(a.hi, b);
  p.lo += a.lo * b;
  return dd_fast_two_sum(p.hi, p.lo);
}

__device__ __forceinline__ DD dd_div_d(double a, DD b) {
  // a / b: quotient, residual a - q b (exact two-product), correction
  double q = a / b.hi;
  DD qb = dd_mul_d(b, q);
  DD r = dd_add({a, 0.}, {-qb.hi, -qb.lo});
  return dd_fast_two_sum(q, r.hi / b.hi);
}

// e^x = 2^k * result, |x.hi| <= 709.79 (double-double argument)
__device__ __noinline__ DD dd_exp(DD x, int& k) {
  double kd = rint(x.hi * CR_INV_LN2);
  k = (int)kd;
  DD r = dd_add(x, dd_add(dd_two_prod(-kd, CR_LN2_HI), dd_two_prod(-kd, CR_LN2_LO)));
  double jd = rint(r.hi * 64.);
  int j = (int)jd;
  j = j < -32 ? -32 : (j > 32 ? 32 : j);
  r = dd_add(r, {-(double)j / 64., 0.});
  DD p = CR_INVFACT[12];
  for (int n = 11; n >= 0; --n) p = dd_add(dd_mul(p, r), CR_INVFACT[n]);
  return dd_mul(p, CR_EXPT[j + 32]);
}

__device__ inline double cr_exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -708.) return exp(x);  // subnormal or zero results
  int k;
  DD e = dd_exp({x, 0.}, k);
  return ldexp(e.hi, k);
}

// l < exp(a) decided as with the correctly rounded exp: the device library's
// exp is within one ulp of e^a, so a comparison outside a band of 2^-50
// relative around it (four ulps and more) is already decided; inside the band
// the correctly rounded value decides.  a >= -708 (normal results).
__device__ __forceinline__ bool lt_exp(double l, double a) {
  double e = exp(a);
  if (l < e * (1. - 0x1p-50)) return true;
  if (l > e * (1. + 0x1p-50)) return false;
  return l < cr_exp(a);
}

// log(x) in double-double, x positive, normal and finite
__device__ __noinline__ DD dd_log(double x) {
  double t = x - 1.;  // exact for x in [1/2, 2]
  if (fabs(t) < 0.03125) {
    DD u = dd_div_d(t, dd_two_sum(2., t));
    DD u2 = dd_mul(u, u);
    DD p = CR_ATANH[CR_NATANH - 1];
    for (int n = CR_NATANH - 2; n >= 0; --n) p = dd_add(dd_mul(p, u2), CR_ATANH[n]);
    return dd_mul(p, u);
  }
  double y0 = log(x);
  int k;
  DD e = dd_exp({-y0, 0.}, k);                 // e^-y0 = e * 2^k
  DD z = dd_add(dd_mul_d(e, ldexp(x, k)), {-1., 0.});  // x e^-y0 - 1, |z| ~ 2^-52
  z = dd_add(z, {-0.5 * z.hi * z.hi, 0.});     // log1p(z) to 2^-150
  return dd_add({y0, 0.}, z);
}

__device__ inline double cr_log(double x) {
  if (!(x >= 2.2250738585072014e-308) || !(x < HUGE_VAL)) return log(x);  // 0, negative, subnormal, inf, nan
  return dd_log(x).hi;
}

__device__ __noinline__ double cr_pow(double x, double y) {
  if (y == 0. || x == 1.) return 1.;
  if (!(x >= 2.2250738585072014e-308) || !(x < HUGE_VAL) || !(fabs(y) < HUGE_VAL)) return pow(x, y);
  DD a = dd_mul_d(dd_log(x), y);
  if (a.hi > 709.782712893384) return HUGE_VAL;
  if (!(a.hi >= -708.)) return pow(x, y);  // subnormal or zero results, nan
  int k;
  DD e = dd_exp(a, k);
  return ldexp(e.hi, k);
}

__device__ __noinline__ double cr_expm1(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -38.) return -1.;  // e^x < 2^-54: -1 + e^x rounds to -1
  if (fabs(x) < 0.03125) {
    DD p = CR_INVFACT[15];
    for (int n = 14; n >= 1; --n) p = dd_add(dd_mul_d(p, x), CR_INVFACT[n]);
    return dd_mul_d(p, x).hi;
  }
  int k;
  DD e = dd_exp({x, 0.}, k);
  DD r = dd_add({ldexp(e.hi, k), ldexp(e.lo, k)}, {-1., 0.});
  return r.hi;
}

}  // namespace cyl
