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

#if defined(IPG_OPCOUNT)
/* The op-counting CPU build prices each of these as the operation it
 * stands for (one division, one transcendental). */
inline double rcp(double x) { return 1.0 / x; }
inline double exp_(double x) { return exp(x); }
inline void sincos_(double x, double &s, double &c) { sincos(x, &s, &c); }
inline void sincospi_(double y, double &s, double &c) { sincos(3.141592653589793 * y, &s, &c); }
#else

IPG_MATH_HD double rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  return (x == 0.) ? __builtin_amdgcn_rcp(x) : y;
#else
  return 1.0 / x;
#endif
}

constexpr double LN2_HI = 0.6931471805599453;
constexpr double LN2_LO = 2.3190468138462996e-17;
constexpr double LOG2E = 1.4426950408889634;

IPG_MATH_HD double exp_(double x)
{
  if (!(fabs(x) < 700.)) return exp(x);
  double k = rint(x * LOG2E);
  double r = fma(-k, LN2_HI, x);
  r = fma(-k, LN2_LO, r);
  double p = 1.0 / 6227020800.0;         /* 1/13! */
  p = fma(p, r, 1.0 / 479001600.0);      /* 1/12! */
  p = fma(p, r, 1.0 / 39916800.0);       /* 1/11! */
  p = fma(p, r, 1.0 / 3628800.0);        /* 1/10! */
  p = fma(p, r, 1.0 / 362880.0);         /* 1/9! */
  p = fma(p, r, 1.0 / 40320.0);          /* 1/8! */
  p = fma(p, r, 1.0 / 5040.0);           /* 1/7! */
  p = fma(p, r, 1.0 / 720.0);            /* 1/6! */
  p = fma(p, r, 1.0 / 120.0);            /* 1/5! */
  p = fma(p, r, 1.0 / 24.0);             /* 1/4! */
  p = fma(p, r, 1.0 / 6.0);              /* 1/3! */
  p = fma(p, r, 0.5);                    /* 1/2! */
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

/* sin r and cos r for |r| <= pi/4 */
IPG_MATH_HD void sincos_reduced(double r, double &s, double &c)
{
  double z = r * r;
  double ps = 1.0 / 355687428096000.0;   /* 1/17! */
  ps = fma(ps, z, -1.0 / 1307674368000.0);
  ps = fma(ps, z, 1.0 / 6227020800.0);
  ps = fma(ps, z, -1.0 / 39916800.0);
  ps = fma(ps, z, 1.0 / 362880.0);
  ps = fma(ps, z, -1.0 / 5040.0);
  ps = fma(ps, z, 1.0 / 120.0);
  ps = fma(ps, z, -1.0 / 6.0);
  s = fma(r * z, ps, r);
  double pc = -1.0 / 6402373705728000.0; /* -1/18! */
  pc = fma(pc, z, 1.0 / 20922789888000.0);
  pc = fma(pc, z, -1.0 / 87178291200.0);
  pc = fma(pc, z, 1.0 / 479001600.0);
  pc = fma(pc, z, -1.0 / 3628800.0);
  pc = fma(pc, z, 1.0 / 40320.0);
  pc = fma(pc, z, -1.0 / 720.0);
  pc = fma(pc, z, 1.0 / 24.0);
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

IPG_MATH_HD void sincos_(double x, double &s, double &c)
{
  if (!(fabs(x) < 1.e5)) {
    s = sin(x);
    c = cos(x);
    return;
  }
  double k = rint(x * TWO_OVER_PI);
  double r = fma(-k, PIO2_HI, x);
  r = fma(-k, PIO2_MID, r);
  r = fma(-k, PIO2_LO, r);
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
  double k = rint(2. * y);
  double t = y - 0.5 * k;
  double r = fma(t, PI_HI, t * PI_MID);
  double sr, cr;
  sincos_reduced(r, sr, cr);
  quadrant((int)k, sr, cr, s, c);
}

#endif

}  // namespace ipg
