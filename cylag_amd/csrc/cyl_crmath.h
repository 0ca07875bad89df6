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
//
// Each function first tries a fast path (Ziv's strategy): a value accurate to
// about 2^-64..2^-69, from short polynomials in double on a table of e^(j/256),
// is returned when its error bound proves how it rounds; otherwise the
// double-double evaluation decides.  The result is the correctly rounded value
// either way; the double-double path runs for a few arguments in a thousand.
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

// ---- fast paths ------------------------------------------------------------

// hi + lo + err rounds to hi for every |err| <= m: (hi, lo) normalized
// (|lo| <= half an ulp of hi), hi normal.  The rounding interval of hi reaches
// half an ulp away from zero and half an ulp of the binade below towards zero
// when |hi| is a power of two; ties are left undecided.
__device__ __forceinline__ bool dd_rounds_to_hi(double hi, double lo, double m) {
  int e;
  const double f = frexp(hi, &e);               // |hi| = |f| 2^e, |f| in [1/2, 1)
  const double half_up = ldexp(1., e - 54);     // half an ulp of hi
  const double half_dn = fabs(f) == 0.5 ? 0.5 * half_up : half_up;
  const double d = (lo != 0. && ((hi > 0.) == (lo > 0.))) ? half_up : half_dn;
  return fabs(lo) + m < d;
}

// e^(xh + xl) = 2^k (hi + lo) with a relative error below 2^-69.4, for
// |xh| <= 709.8 and |xl| <= 2^-43.  Reduction: kd ln2_hi = p + pe exactly
// (two-product), xh - p is exact (Sterbenz), and the small terms xl, pe,
// kd ln2_lo add up with an error below 2^-94; then r = j/256 + s, with
// r.hi - j/256 exact (Sterbenz) and |s| <= 2^-9.
// e^s = 1 + s.hi + q(s.hi) + s.lo (1 + s.hi), q(t) = t^2/2 + ... + t^7/7!
// (truncation 2^-87, rounding of q below 2^-70.3 since |q| <= 2^-19), times
// the double-double e^(j/256) (2^-107).
__device__ __forceinline__ DD exp_fast(double xh, double xl, int& k) {
  const double kd = rint(xh * CR_INV_LN2);
  k = (int)kd;
  const double p = kd * CR_LN2_HI, pe = fma(kd, CR_LN2_HI, -p);
  const double c = (xl - pe) - kd * CR_LN2_LO;
  const DD r = dd_two_sum(xh - p, c);
  const double jd = rint(r.hi * 256.);
  int j = (int)jd;
  j = j < -CR_NT256 ? -CR_NT256 : (j > CR_NT256 ? CR_NT256 : j);
  const DD s = dd_two_sum(r.hi - (double)j * (1. / 256.), r.lo);
  const double t = s.hi;
  const double q = (t * t) * (0.5 + t * (1. / 6. + t * (1. / 24. + t * (1. / 120. + t * (1. / 720. + t * (1. / 5040.))))));
  const double tail = (s.lo + t * s.lo) + q;
  DD E = dd_two_sum(1., t);
  E = dd_fast_two_sum(E.hi, E.lo + tail);
  return dd_mul(E, CR_EXPT256[j + CR_NT256]);
}

// log(x) in double-double (x positive, normal, finite); err: a bound on its
// absolute error, at most 2^-68.9.  |x - 1| < 1/32: 2 atanh(u), u = t / (2 + t)
// in double-double, 2u (1 + P(u^2)) with P = u^2/3 + ... + u^10/11 in double
// from u.hi (|P| <= 2^-13.6; its rounding 2^-64.6, the products 2^-66.6, the
// neglected u.lo terms 2^-65.6 and 2^-66.6, truncation 2^-75.6: 2^-63.9
// relative, bounded by 2^-63.5 |log x|).  Otherwise y0 = log(x) from the
// library (within an ulp) and the correction z = x e^-y0 - 1 (|z| ~ 2^-52,
// absolute error 2^-69.4 from the fast exp), log x = y0 + z - z^2/2: absolute
// error 2^-69.4, bounded by 2^-68.5.
__device__ __forceinline__ DD log_fast(double x, double& err) {
  const double t = x - 1.;  // exact for x in [1/2, 2]
  if (fabs(t) < 0.03125) {
    const DD u = dd_div_d(t, dd_two_sum(2., t));
    const double w = u.hi * u.hi;
    const double P = w * (1. / 3. + w * (1. / 5. + w * (1. / 7. + w * (1. / 9. + w * (1. / 11.)))));
    const DD u2 = {2. * u.hi, 2. * u.lo};
    const DD l = dd_fast_two_sum(u2.hi, u2.lo + u2.hi * P);
    err = fabs(l.hi) * 0x1.6a09e667f3bcdp-64;  // 2^-63.5 |log x|
    return l;
  }
  const double y0 = log(x);
  int k;
  const DD e = exp_fast(-y0, 0., k);
  DD z = dd_add(dd_mul_d(e, ldexp(x, k)), {-1., 0.});
  z.lo -= 0.5 * z.hi * z.hi;
  err = 0x1.6a09e667f3bcdp-69;  // 2^-68.5
  return dd_add({y0, 0.}, z);
}

__device__ __forceinline__ bool cr_exp_try(double x, double& out) {
  int k;
  const DD e = exp_fast(x, 0., k);
  out = ldexp(e.hi, k);
  return dd_rounds_to_hi(e.hi, e.lo, fabs(e.hi) * 0x1p-68);
}

// |x| < 1/32: x + x^2/2 + x^3 R(x), x^2 by two-product, R = 1/6 + ... + x^8/11!
// in double (relative error of the sum below 2^-63.3); otherwise e^x - 1 in
// double-double from the fast exp (relative error at most 33 x 2^-69.4).
__device__ __forceinline__ bool cr_expm1_try(double x, double& out) {
  DD r;
  double m;
  if (fabs(x) < 0.03125) {
    const DD x2 = dd_two_prod(x, x);
    const double R = 1. / 6. + x * (1. / 24. + x * (1. / 120. + x * (1. / 720. + x * (1. / 5040. + x * (1. / 40320. +
                     x * (1. / 362880. + x * (1. / 3628800. + x * (1. / 39916800.))))))));
    const double rest = (x * x2.hi) * R;
    const DD a = dd_two_sum(x, 0.5 * x2.hi);
    r = dd_fast_two_sum(a.hi, a.lo + (0.5 * x2.lo + rest));
    m = fabs(r.hi) * 0x1.6a09e667f3bcdp-63;
  } else {
    int k;
    const DD e = exp_fast(x, 0., k);
    r = dd_add({ldexp(e.hi, k), ldexp(e.lo, k)}, {-1., 0.});
    m = fabs(r.hi) * 0x1.6a09e667f3bcdp-64;
  }
  out = r.hi;
  return dd_rounds_to_hi(r.hi, r.lo, m);
}

// ---- correctly rounded functions ------------------------------------------

__device__ __noinline__ double cr_exp_dd(double x) {
  int k;
  DD e = dd_exp({x, 0.}, k);
  return ldexp(e.hi, k);
}

__device__ __forceinline__ double cr_exp(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -708.) return exp(x);  // subnormal or zero results
  double r;
  if (cr_exp_try(x, r)) return r;
  return cr_exp_dd(x);
}

// l < exp(a) decided as with the correctly rounded exp: the device library's
// exp is within one ulp of e^a, so a comparison outside a band of 2^-50
// relative around it (four ulps and more) is already decided; inside the band
// the correctly rounded value decides.  a >= -708 (normal results).
__device__ __noinline__ double cr_exp_call(double x) { return cr_exp(x); }

__device__ __forceinline__ bool lt_exp(double l, double a) {
  double e = exp(a);
  if (l < e * (1. - 0x1p-50)) return true;
  if (l > e * (1. + 0x1p-50)) return false;
  return l < cr_exp_call(a);  // rare: out of line, so callers keep their registers
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

__device__ __forceinline__ bool cr_log_try(double x, double& out) {
  double err;
  const DD l = log_fast(x, err);
  out = l.hi;
  return dd_rounds_to_hi(l.hi, l.lo, err);
}

__device__ __noinline__ double cr_log_dd(double x) { return dd_log(x).hi; }

__device__ __forceinline__ double cr_log(double x) {
  if (!(x >= 2.2250738585072014e-308) || !(x < HUGE_VAL)) return log(x);  // 0, negative, subnormal, inf, nan
  if (x == 1.) return 0.;
  double r;
  if (cr_log_try(x, r)) return r;
  return cr_log_dd(x);
}

// cr_log out of line, for rarely taken paths
__device__ __noinline__ double cr_log_call(double x) { return cr_log(x); }

// pow = e^(y log x): the fast log's absolute error err (at most 2^-68.9)
// becomes |y| err in a = y log x, with the product's rounding (2^-104 |a|),
// a relative error of e^a on top of the fast exp's 2^-69.4 (bounded by 2^-68).
// x positive, normal, finite, x != 1, y finite and nonzero, -708 <= a <= 709.78.
__device__ __forceinline__ bool cr_pow_try(double x, double y, double& out) {
  double err;
  const DD a = dd_mul_d(log_fast(x, err), y);
  if (!(a.hi >= -708.) || a.hi > 709.782712893384) return false;
  int k;
  const DD e = exp_fast(a.hi, a.lo, k);
  out = ldexp(e.hi, k);
  return dd_rounds_to_hi(e.hi, e.lo, fabs(e.hi) * (fabs(y) * err + fabs(a.hi) * 0x1p-100 + 0x1p-68));
}

__device__ __noinline__ double cr_pow_dd(double x, double y) {
  DD a = dd_mul_d(dd_log(x), y);
  if (a.hi > 709.782712893384) return HUGE_VAL;
  if (!(a.hi >= -708.)) return pow(x, y);  // subnormal or zero results, nan
  int k;
  DD e = dd_exp(a, k);
  return ldexp(e.hi, k);
}

__device__ __forceinline__ double cr_pow(double x, double y) {
  if (y == 0. || x == 1.) return 1.;
  if (!(x >= 2.2250738585072014e-308) || !(x < HUGE_VAL) || !(fabs(y) < HUGE_VAL)) return pow(x, y);
  double r;
  if (cr_pow_try(x, y, r)) return r;
  return cr_pow_dd(x, y);
}

__device__ __noinline__ double cr_expm1_dd(double x);

__device__ __forceinline__ double cr_expm1(double x) {
  if (!(x == x)) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -38.) return -1.;  // e^x < 2^-54: -1 + e^x rounds to -1
  double r;
  if (cr_expm1_try(x, r)) return r;
  return cr_expm1_dd(x);
}

// the double-double expm1 (|x| <= 709.78, x >= -38)
__device__ __noinline__ double cr_expm1_dd(double x) {
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
