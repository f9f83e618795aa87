/*
 * Elementary functions for the per-step loops of the ray pipeline.
 *
 * A gfx950 wave issues one FP64 instruction per 4 cycles, and the geodesic
 * and transfer recurrences of one ray run on one lane, so the instruction
 * count of a step is its latency.  The general-purpose library routines
 * cover arguments and special cases these loops never see (huge arguments,
 * IEEE-exact division); these versions cover the ranges that occur and
 * fall back to the library routine outside them.
 *
 *   rcp(x)       1/x: the hardware reciprocal estimate refined by two Newton
 *                steps (e -> e^2 each), on the device; 1/x on the host.
 *   exp_(x)      Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln 2
 *                split in two parts, the Taylor series of e^r to r^13
 *                (truncation < 4e-18), and scaling by 2^k.
 *
 * The series are evaluated in Estrin's scheme (terms in pairs, then
 * combined with the argument's powers r^2, r^4, r^8), whose dependency
 * depth is about the logarithm of the term count: a dependent FP64 FMA
 * issues every ~6 cycles on gfx950 against ~4 for independent ones, and
 * the geodesic, chain and batch-transfer lanes run one wave per SIMD.
 *   sincos_(x)   reduction x = k pi/2 + r with pi/2 in three parts (FMA),
 *                |r| <= pi/4, Taylor series of sin to r^17 and cos to r^18
 *                (truncation < 1e-19), quadrant select by k mod 4.
 *   sincospi_(y) sin(pi y), cos(pi y): y = k/2 + t with t exact, r = pi t.
 *
 * The series coefficients are 1/n!, written as such; the split constants
 * are pi/2, pi and ln 2 rounded in successive pieces.  Accuracy: a few ulp
 * over the covered ranges (checked against the host's libm and mpmath).
 */
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define IPG_MATH_HD __host__ __device__ inline
#else
#define IPG_MATH_HD inline
#endif

namespace ipg {

/* The library routines the fast paths below fall back to outside their
 * ranges, called out of line on the device: each fast path's rare branch is
 * then a call instead of an inlined copy of the routine (the packet and
 * batch-transfer kernels inline dozens of call sites; their code shrinks by
 * about a third, which keeps more of it in the instruction cache). */
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline)) double lib_exp(double x) { return exp(x); }
__device__ __attribute__((noinline)) double lib_log(double x) { return log(x); }
__device__ __attribute__((noinline)) double lib_pow(double a, double b) { return pow(a, b); }
__device__ __attribute__((noinline)) double lib_cbrt(double x) { return cbrt(x); }
__device__ __attribute__((noinline)) void lib_sincos(double x, double *s, double *c)
{
  *s = sin(x);
  *c = cos(x);
}
#else
inline double lib_exp(double x) { return exp(x); }
inline double lib_log(double x) { return log(x); }
inline double lib_pow(double a, double b) { return pow(a, b); }
inline double lib_cbrt(double x) { return cbrt(x); }
inline void lib_sincos(double x, double *s, double *c)
{
  *s = sin(x);
  *c = cos(x);
}
#endif

#if !defined(IPG_OPCOUNT)
typedef double ipg_f64; /* plain double for compile-time constant tables */
#endif

#if defined(IPG_OPCOUNT)
/* The op-counting CPU build prices each of these as the operation it
 * stands for (one division, one transcendental). */
inline double rcp(double x) { return 1.0 / x; }
inline double exp_(double x) { return exp(x); }
inline void sincos_(double x, double &s, double &c) { sincos(x, &s, &c); }
inline void sincospi_(double y, double &s, double &c) { sincos(3.141592653589793 * y, &s, &c); }
inline double exp_r(double x) { return exp(x); }
inline double exp_neg(double x) { return exp(x); }
inline void sincos_r(double x, double &s, double &c) { sincos(x, &s, &c); }
inline void sincospi_r(double y, double &s, double &c) { sincos(3.141592653589793 * y, &s, &c); }
inline double log_(double x) { return log(x); }
inline double pow_(double a, double b) { return pow(a, b); }
inline double cbrt_(double x) { return cbrt(x); }
#else

IPG_MATH_HD double rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  /* x = 0: the hardware value (inf) without the refinement, which would
   * turn it into NaN.  Callers never pass infinities. */
  return (x == 0.) ? __builtin_amdgcn_rcp(x) : y;
#else
  return 1.0 / x;
#endif
}

constexpr double LN2_HI = 0.6931471805599453;
constexpr double LN2_LO = 2.3190468138462996e-17;
constexpr double LOG2E = 1.4426950408889634;

/* exp_r, sincos_r, sincospi_r: the arithmetic of exp_, sincos_ and
 * sincospi_ without the library fallback, for callers whose arguments are
 * finite and in range (|x| < 700, |x| < 1e5, |y| < 1e9; NaN propagates).
 * Their bodies have no branch, so that the compiler can interleave two
 * independent evaluations (the geodesic step evaluates the metric at two
 * points at once). */
IPG_MATH_HD double exp_r(double x)
{
  double k = rint(x * LOG2E);
  double r = fma(-k, LN2_HI, x);
  r = fma(-k, LN2_LO, r);
  /* q = sum_{n=2}^{13} r^(n-2)/n! in Estrin's scheme (pairs in r, then
   * in r^2, r^4, r^8: a dependency depth of 5 instead of 12, for the
   * lanes that run one wave per SIMD and wait on each dependent FMA), then
   * p = 1 + r (1 + r q) as the last two steps of Horner's rule */
  const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  const double a0 = fma(1.0 / 6.0, r, 0.5);                       /* 1/2!, 1/3! */
  const double a1 = fma(1.0 / 120.0, r, 1.0 / 24.0);              /* 1/4!, 1/5! */
  const double a2 = fma(1.0 / 5040.0, r, 1.0 / 720.0);            /* 1/6!, 1/7! */
  const double a3 = fma(1.0 / 362880.0, r, 1.0 / 40320.0);        /* 1/8!, 1/9! */
  const double a4 = fma(1.0 / 39916800.0, r, 1.0 / 3628800.0);    /* 1/10!, 1/11! */
  const double a5 = fma(1.0 / 6227020800.0, r, 1.0 / 479001600.0);/* 1/12!, 1/13! */
  const double b0 = fma(a1, r2, a0), b1 = fma(a3, r2, a2), b2 = fma(a5, r2, a4);
  const double q = fma(b2, r8, fma(b1, r4, b0));
  const double p = fma(fma(q, r, 1.0), r, 1.0);
  return ldexp(p, (int)k);
}

IPG_MATH_HD double exp_(double x)
{
  if (!(fabs(x) < 700.)) return lib_exp(x);
  return exp_r(x);
}

/* e^x for x <= 0 without a branch: arguments below -800 (whose e^x is 0)
 * are clamped there, NaN propagates; the same value as exp_ on [-700, 0]. */
IPG_MATH_HD double exp_neg(double x)
{
  return exp_r(x != x ? x : fmax(x, -800.));
}

/* sin r and cos r for |r| <= pi/4 */
IPG_MATH_HD void sincos_reduced(double r, double &s, double &c)
{
  /* both series in z = r^2 in Estrin's scheme (pairs, then z^2, z^4) */
  const double z = r * r, z2 = z * z, z4 = z2 * z2;
  const double s0 = fma(1.0 / 120.0, z, -1.0 / 6.0);                          /* -1/3!, 1/5! */
  const double s1 = fma(1.0 / 362880.0, z, -1.0 / 5040.0);                    /* -1/7!, 1/9! */
  const double s2 = fma(1.0 / 6227020800.0, z, -1.0 / 39916800.0);            /* -1/11!, 1/13! */
  const double s3 = fma(1.0 / 355687428096000.0, z, -1.0 / 1307674368000.0);  /* -1/15!, 1/17! */
  const double ps = fma(fma(s3, z2, s2), z4, fma(s1, z2, s0));
  s = fma(r * z, ps, r);
  const double c0 = fma(-1.0 / 720.0, z, 1.0 / 24.0);                         /* 1/4!, -1/6! */
  const double c1 = fma(-1.0 / 3628800.0, z, 1.0 / 40320.0);                  /* 1/8!, -1/10! */
  const double c2 = fma(-1.0 / 87178291200.0, z, 1.0 / 479001600.0);          /* 1/12!, -1/14! */
  const double c3 = fma(-1.0 / 6402373705728000.0, z, 1.0 / 20922789888000.0);/* 1/16!, -1/18! */
  const double pc = fma(fma(c3, z2, c2), z4, fma(c1, z2, c0));
  c = fma(z, fma(z, pc, -0.5), 1.0);
}

IPG_MATH_HD void quadrant(int q, double sr, double cr, double &s, double &c)
{
  q &= 3;
  double a = (q & 1) ? cr : sr;   /* sin: S, C, -S, -C */
  double b = (q & 1) ? sr : cr;   /* cos: C, -S, -C, S */
  s = (q & 2) ? -a : a;
  c = ((q + 1) & 2) ? -b : b;
}

constexpr double PIO2_HI = 1.5707963267948966;
constexpr double PIO2_MID = 6.123233995736766e-17;
constexpr double PIO2_LO = -1.4973849048591698e-33;
constexpr double TWO_OVER_PI = 0.6366197723675814;
constexpr double PI_HI = 3.141592653589793;
constexpr double PI_MID = 1.2246467991473532e-16;

IPG_MATH_HD void sincos_r(double x, double &s, double &c)
{
  double k = rint(x * TWO_OVER_PI);
  double r = fma(-k, PIO2_HI, x);
  r = fma(-k, PIO2_MID, r);
  r = fma(-k, PIO2_LO, r);
  double sr, cr;
  sincos_reduced(r, sr, cr);
  quadrant((int)k, sr, cr, s, c);
}

IPG_MATH_HD void sincos_(double x, double &s, double &c)
{
  if (!(fabs(x) < 1.e5)) {
    lib_sincos(x, &s, &c);
    return;
  }
  sincos_r(x, s, c);
}

IPG_MATH_HD void sincospi_r(double y, double &s, double &c)
{
  double k = rint(2. * y);
  double t = y - 0.5 * k;
  double r = fma(t, PI_HI, t * PI_MID);
  double sr, cr;
  sincos_reduced(r, sr, cr);
  quadrant((int)k, sr, cr, s, c);
}

IPG_MATH_HD void sincospi_(double y, double &s, double &c)
{
  if (!(fabs(y) < 1.e9)) {
    s = sin(PI_HI * y);
    c = cos(PI_HI * y);
    return;
  }
  sincospi_r(y, s, c);
}

constexpr double SQRT_HALF = 0.7071067811865476;

/* log x: x = 2^e m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
 * s = (m - 1)/(m + 1), |s| <= 0.1716, series to s^23 (truncation < 1e-18
 * relative); outside (1e-300, 1e300) and for 0, negatives, inf, NaN the
 * library routine. */
IPG_MATH_HD double log_(double x)
{
  if (!(x >= 1.e-300 && x <= 1.e300)) return lib_log(x);
  int e;
  double m = frexp(x, &e);
  if (m < SQRT_HALF) {
    m *= 2.;
    e -= 1;
  }
  double s = (m - 1.) * rcp(m + 1.);
  /* sum_{n=0}^{10} z^n / (2n + 3) in Estrin's scheme (pairs, then z^2,
   * z^4, z^8) */
  const double z = s * s, z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
  const double a0 = fma(1.0 / 5.0, z, 1.0 / 3.0), a1 = fma(1.0 / 9.0, z, 1.0 / 7.0);
  const double a2 = fma(1.0 / 13.0, z, 1.0 / 11.0), a3 = fma(1.0 / 17.0, z, 1.0 / 15.0);
  const double a4 = fma(1.0 / 21.0, z, 1.0 / 19.0);
  const double b0 = fma(a1, z2, a0), b1 = fma(a3, z2, a2), b2 = fma(1.0 / 23.0, z2, a4);
  const double p = fma(b2, z8, fma(b1, z4, b0));
  double at2 = 2. * fma(s * z, p, s);
  double de = (double)e;
  return fma(de, LN2_HI, fma(de, LN2_LO, at2));
}

/* a^b for a > 0 as e^(b log a); a = 0 gives 0 for b > 0 as pow does. */
IPG_MATH_HD double pow_(double a, double b)
{
  if (!(a > 0.)) return lib_pow(a, b);
  return exp_(b * log_(a));
}

/* cube root of x > 0 (library routine otherwise) */
IPG_MATH_HD double cbrt_(double x)
{
  if (!(x > 0.)) return lib_cbrt(x);
  double y = exp_(log_(x) * (1.0 / 3.0));
  /* one Newton step on y^3 = x restores the last bits lost to the
   * log/exp round trip */
  return y - (y * y * y - x) * rcp(3. * y * y);
}

#endif

}  // namespace ipg
