/*
 * Device physics of the per-pixel ray pipeline (gfx950, FP64).
 *
 * Everything a lane needs to trace one ray and integrate polarized transfer
 * along it.  The formulation is built for a GPU lane, not translated from the
 * reference loops:
 *   - the metric of KS/MKS/FMKS/eMinkowski is evaluated once per point with
 *     its exact X1/X2 derivatives (chain rule through r(X1), theta(X1,X2)) and
 *     its closed-form inverse; the reference instead differences 8 metric
 *     evaluations (DEL = 1e-7) and inverts by LU (src/geometry.c:93-144).
 *     The two agree to the reference's own finite-difference round-off
 *     (~1e-9 relative in Gamma).  t and phi derivatives vanish identically
 *     for these stationary, axisymmetric metrics and are never formed.
 *   - Gamma is never materialized: the geodesic needs Gamma^i_jk K^j K^k and
 *     parallel transport needs M^i_k = Gamma^i_kl K^l; both are contracted
 *     directly from dg/dX1, dg/dX2 and g^-1.
 *   - the coherency tensor N is Hermitian and is kept as its upper triangle
 *     (4 real diagonal + 6 complex entries, 16 doubles); the reference keeps
 *     the full complex 4x4 (src/ipolarray.c:253-289, :710-748).
 *   - the Stokes update in the plasma frame only touches the [1..2][1..2]
 *     tetrad block (src/ipolarray.c:673-682); it is applied to N as
 *     N += Econ^T dT Econ instead of a full rank-2 round trip.
 */
#pragma once

#include <math.h>
#include "ipgpu.h"
#include "ipg_bessel_coeffs.h"
#include "ipg_math.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define IPG_HD __host__ __device__ inline
/* for the per-step bodies whose outputs must stay in registers */
#define IPG_HD_INLINE __host__ __device__ inline __attribute__((always_inline))
#else
#define IPG_HD inline
#define IPG_HD_INLINE inline
#endif

namespace ipg {

/* constants of src/constants.h */
constexpr double EE = 4.80320680e-10;
constexpr double CL = 2.99792458e10;
constexpr double ME = 9.1093826e-28;
constexpr double MP = 1.67262171e-24;
constexpr double HPL = 6.6260693e-27;
constexpr double SMALL = 1.e-40;
constexpr double PI = 3.14159265358979323846;

/* packed index of a symmetric 4x4 (i <= j): 00 01 02 03 11 12 13 22 23 33 */
IPG_HD constexpr int sy(int i, int j)
{
  return (i <= j) ? (i * 4 - (i * (i - 1)) / 2 + (j - i)) : (j * 4 - (j * (j - 1)) / 2 + (i - j));
}

IPG_HD double ipow_d(double x, int n)
{
  double r = 1.0, b = x;
  unsigned u = (n < 0) ? (unsigned)(-n) : (unsigned)n;
  while (u) {
    if (u & 1u) r *= b;
    b *= b;
    u >>= 1;
  }
  return (n < 0) ? 1.0 / r : r;
}

/* x^alpha for FMKS's poly_alpha; integer-valued exponents (iharm3d uses 14)
 * by squaring, which is also defined for negative x as pow() is there. */
IPG_HD double powa(double x, double alpha)
{
  double ai = floor(alpha);
  if (ai == alpha && fabs(alpha) < 64.0) return ipow_d(x, (int)ai);
  return pow(x, alpha);
}

/* x^n for 0 <= n < 64: the products of ipow_d in the same order, selected
 * instead of branched on. */
IPG_HD_INLINE double ipow_sel(double x, int n)
{
  double r = 1.0, b = x;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    r = ((n >> i) & 1) ? r * b : r;
    b = b * b;
  }
  return r;
}

/* Metric kinds with straight-line evaluation code (template argument of
 * theta_t / metric_t): eMinkowski, MKS, FMKS with an integer poly_alpha - 1
 * in [0, 64) (x^n by ipow_sel), other FMKS (powa). */
enum { MK_MINK = 0, MK_MKS = 1, MK_FMKS_INT = 2, MK_FMKS = 3, MK_ANY = 4 };

IPG_HD int metric_kind(const ipgpu_geom &G)
{
  if (G.metric == IPGPU_METRIC_EMINKOWSKI) return MK_MINK;
  if (G.metric == IPGPU_METRIC_MKS) return MK_MKS;
  double a1 = G.poly_alpha - 1.;
  return (floor(a1) == a1 && a1 >= 0. && a1 < 64.) ? MK_FMKS_INT : MK_FMKS;
}

/* ------------------------------------------------------------------ metric */

struct Metric {
  double r, th;
  double g[10];   /* g_{mu nu} */
  double gc[10];  /* g^{mu nu} */
  double d1[10];  /* d g_{mu nu} / dX1 */
  double d2[10];  /* d g_{mu nu} / dX2 */
  double sqrtmg;  /* sqrt(|det g|) */
};

struct Theta {
  double th, t1, t2, t11, t12, t22;
};

/* theta(X1, X2) and derivatives for MKS/FMKS: src/coordinates.c:69-81
 * (values) and :278-305 (first derivatives; the second derivatives follow
 * from the same closed form).  X2 in [0, 1] and X1 within the grid keep the
 * in-range math in range. */
template <int MK>
IPG_HD_INLINE void theta_t(const ipgpu_geom &G, double X1, double X2, Theta &T)
{
  double s2p, c2p;
  sincospi_r(2. * X2, s2p, c2p);
  double hs = G.hslope;
  double thG = PI * X2 + ((1. - hs) / 2.) * s2p;
  double dG = PI + (1. - hs) * PI * c2p;
  double ddG = -2. * PI * PI * (1. - hs) * s2p;
  if constexpr (MK == MK_MKS || MK == MK_MINK) {
    T.th = thG;
    T.t1 = 0.;
    T.t2 = dG;
    T.t11 = 0.;
    T.t12 = 0.;
    T.t22 = ddG;
    return;
  } else {
    double y = 2. * X2 - 1.;
    double al = G.poly_alpha, xt = G.poly_xt, pn = G.poly_norm;
    double ixt = rcp(xt);
    double P = y * ixt;
    double Pa1 = (MK == MK_FMKS_INT) ? ipow_sel(P, (int)(al - 1.)) : powa(P, al - 1.);
    double Pa = Pa1 * P;
    double thJ = pn * y * (1. + Pa * rcp(al + 1.)) + 0.5 * PI;
    double dJ = 2. * pn * (1. + Pa);
    double ddJ = 4. * pn * al * Pa1 * ixt;
    double E = exp_r(G.mks_smooth * (G.startx[1] - X1));
    double ms = G.mks_smooth;
    T.th = thG + E * (thJ - thG);
    T.t1 = -ms * E * (thJ - thG);
    T.t2 = dG + E * (dJ - dG);
    T.t11 = ms * ms * E * (thJ - thG);
    T.t12 = -ms * E * (dJ - dG);
    T.t22 = ddG + E * (ddJ - ddG);
  }
}

/* r, theta only (src/coordinates.c:46-97) */
IPG_HD void bl_coord(const ipgpu_geom &G, double X1, double X2, double &r, double &th)
{
  r = exp_(X1);
  if (G.metric == IPGPU_METRIC_EMINKOWSKI) {
    th = X2;
    return;
  }
  double s2p, c2p;
  sincospi_(2. * X2, s2p, c2p);
  double thG = PI * X2 + ((1. - G.hslope) / 2.) * s2p;
  if (G.metric == IPGPU_METRIC_MKS) {
    th = thG;
    return;
  }
  double y = 2. * X2 - 1.;
  double thJ = G.poly_norm * y * (1. + powa(y * rcp(G.poly_xt), G.poly_alpha) * rcp(G.poly_alpha + 1.))
      + 0.5 * PI;
  th = thG + exp_(G.mks_smooth * (G.startx[1] - X1)) * (thJ - thG);
}

/* Metric, inverse and X1/X2 derivatives at (X1, X2).  Same tensor as
 * gcov_func (src/coordinates.c:175-213, gcov_ks :216-239, set_dxdX
 * :263-348): g_MKS = J^T g_KS J with J = dx^KS/dX.  One straight-line body
 * per metric kind (MK_*). */
template <int MK>
IPG_HD_INLINE void metric_t(const ipgpu_geom &G, double X1, double X2, Metric &m)
{
  double r = exp_r(X1);
  m.r = r;
  if constexpr (MK == MK_MINK) {
    double th = X2, s, c;
    sincos_r(th, s, c);
    double r2 = r * r;
    m.th = th;
    for (int n = 0; n < 10; n++) {
      m.g[n] = 0.;
      m.gc[n] = 0.;
      m.d1[n] = 0.;
      m.d2[n] = 0.;
    }
    m.g[0] = -1.;
    m.g[4] = r2;
    m.g[7] = r2;
    m.g[9] = r2 * s * s;
    m.gc[0] = -1.;
    double ir2 = rcp(r2);
    m.gc[4] = ir2;
    m.gc[7] = ir2;
    m.gc[9] = rcp(r2 * s * s);
    m.d1[4] = 2. * r2;
    m.d1[7] = 2. * r2;
    m.d1[9] = 2. * r2 * s * s;
    m.d2[9] = 2. * r2 * s * c;
    m.sqrtmg = r2 * r * fabs(s);
    return;
  } else {
  Theta T;
  theta_t<MK>(G, X1, X2, T);
  m.th = T.th;
  double a = G.a, a2 = a * a;
  double s, c;
  sincos_r(T.th, s, c);
  double s2 = s * s, s2t = 2. * s * c;
  double Sig = r * r + a2 * c * c;
  double Sr = 2. * r, St = -2. * a2 * c * s;
  double iS = rcp(Sig);
  double q = 2. * r * iS;
  double qr = (2. - q * Sr) * iS;
  double qt = -q * St * iS;
  double opq = 1. + q;
  /* KS components and (r, theta) partials */
  double k00 = -1. + q, k01 = q, k03 = -a * s2 * q, k11 = opq, k13 = -a * s2 * opq, k22 = Sig;
  double k33 = s2 * (Sig + a2 * s2 * opq);
  double k03r = -a * s2 * qr, k03t = -a * (s2t * q + s2 * qt);
  double k13r = -a * s2 * qr, k13t = -a * (s2t * opq + s2 * qt);
  double k33r = s2 * Sr + a2 * s2 * s2 * qr;
  double k33t = s2t * Sig + s2 * St + a2 * (2. * s2 * s2t * opq + s2 * s2 * qt);
  /* chain rule: d/dX1 = r d/dr + t1 d/dth, d/dX2 = t2 d/dth */
  double t1 = T.t1, t2 = T.t2;
  double k00_1 = qr * r + qt * t1, k00_2 = qt * t2;
  double k03_1 = k03r * r + k03t * t1, k03_2 = k03t * t2;
  double k13_1 = k13r * r + k13t * t1, k13_2 = k13t * t2;
  double k22_1 = Sr * r + St * t1, k22_2 = St * t2;
  double k33_1 = k33r * r + k33t * t1, k33_2 = k33t * t2;
  /* MKS components: J11 = r, J21 = t1, J22 = t2 */
  m.g[0] = k00;
  m.g[1] = r * k01;
  m.g[2] = 0.;
  m.g[3] = k03;
  m.g[4] = r * r * k11 + t1 * t1 * k22;
  m.g[5] = t1 * t2 * k22;
  m.g[6] = r * k13;
  m.g[7] = t2 * t2 * k22;
  m.g[8] = 0.;
  m.g[9] = k33;
  /* X1 derivatives: dr/dX1 = r, dt1/dX1 = t11, dt2/dX1 = t12 */
  m.d1[0] = k00_1;
  m.d1[1] = r * k01 + r * k00_1;
  m.d1[2] = 0.;
  m.d1[3] = k03_1;
  m.d1[4] = 2. * r * r * k11 + r * r * k00_1 + 2. * t1 * T.t11 * k22 + t1 * t1 * k22_1;
  m.d1[5] = (T.t11 * t2 + t1 * T.t12) * k22 + t1 * t2 * k22_1;
  m.d1[6] = r * k13 + r * k13_1;
  m.d1[7] = 2. * t2 * T.t12 * k22 + t2 * t2 * k22_1;
  m.d1[8] = 0.;
  m.d1[9] = k33_1;
  /* X2 derivatives: dr/dX2 = 0, dt1/dX2 = t12, dt2/dX2 = t22 */
  m.d2[0] = k00_2;
  m.d2[1] = r * k00_2;
  m.d2[2] = 0.;
  m.d2[3] = k03_2;
  m.d2[4] = r * r * k00_2 + 2. * t1 * T.t12 * k22 + t1 * t1 * k22_2;
  m.d2[5] = (T.t12 * t2 + t1 * T.t22) * k22 + t1 * t2 * k22_2;
  m.d2[6] = r * k13_2;
  m.d2[7] = 2. * t2 * T.t22 * k22 + t2 * t2 * k22_2;
  m.d2[8] = 0.;
  m.d2[9] = k33_2;
  /* inverse: KS contravariant metric, then g^MKS = Jinv g^KS Jinv^T with
   * Jinv11 = 1/r, Jinv21 = -t1/(r t2), Jinv22 = 1/t2 */
  double Del = r * r - 2. * r + a2;
  double c00 = -opq, c01 = q, c11 = Del * iS, c13 = a * iS, c22 = iS, c33 = iS * rcp(s2);
  double ir = rcp(r), w = rcp(t2), u = -t1 * ir * w;
  m.gc[0] = c00;
  m.gc[1] = ir * c01;
  m.gc[2] = u * c01;
  m.gc[3] = 0.;
  m.gc[4] = ir * ir * c11;
  m.gc[5] = ir * u * c11;
  m.gc[6] = ir * c13;
  m.gc[7] = u * u * c11 + w * w * c22;
  m.gc[8] = u * c13;
  m.gc[9] = c33;
  m.sqrtmg = Sig * fabs(s) * r * fabs(t2);
  }
}

IPG_HD_INLINE void metric(const ipgpu_geom &G, double X1, double X2, Metric &m)
{
  switch (metric_kind(G)) {
  case MK_MINK: metric_t<MK_MINK>(G, X1, X2, m); break;
  case MK_MKS: metric_t<MK_MKS>(G, X1, X2, m); break;
  case MK_FMKS_INT: metric_t<MK_FMKS_INT>(G, X1, X2, m); break;
  default: metric_t<MK_FMKS>(G, X1, X2, m); break;
  }
}

/* metric_t for a kind fixed at compile time, metric() for MK_ANY */
template <int MK>
IPG_HD_INLINE void metric_k(const ipgpu_geom &G, double X1, double X2, Metric &m)
{
  if constexpr (MK == MK_ANY) metric(G, X1, X2, m);
  else metric_t<MK>(G, X1, X2, m);
}

IPG_HD void lower(const double g[10], const double u[4], double l[4])
{
  l[0] = g[0] * u[0] + g[1] * u[1] + g[2] * u[2] + g[3] * u[3];
  l[1] = g[1] * u[0] + g[4] * u[1] + g[5] * u[2] + g[6] * u[3];
  l[2] = g[2] * u[0] + g[5] * u[1] + g[7] * u[2] + g[8] * u[3];
  l[3] = g[3] * u[0] + g[6] * u[1] + g[8] * u[2] + g[9] * u[3];
}

IPG_HD double dot4(const double a[4], const double b[4])
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

/* Index lowering with the metrics metric_t forms (Kerr-Schild-based MKS and
 * FMKS, and Minkowski), which have g_02 = g_23 = 0 (so have their X1, X2
 * derivatives) and g^03 = 0: the products with these zeros are left out
 * (the compiler may not drop a product with a zero under IEEE rules; the sums
 * are those of lower() up to the sign of an exact zero). */
IPG_HD void lower_g(const double g[10], const double u[4], double l[4])
{
  l[0] = g[0] * u[0] + g[1] * u[1] + g[3] * u[3];
  l[1] = g[1] * u[0] + g[4] * u[1] + g[5] * u[2] + g[6] * u[3];
  l[2] = g[5] * u[1] + g[7] * u[2];
  l[3] = g[3] * u[0] + g[6] * u[1] + g[9] * u[3];
}

IPG_HD void raise_gc(const double gc[10], const double u[4], double l[4])
{
  l[0] = gc[0] * u[0] + gc[1] * u[1] + gc[2] * u[2];
  l[1] = gc[1] * u[0] + gc[4] * u[1] + gc[5] * u[2] + gc[6] * u[3];
  l[2] = gc[2] * u[0] + gc[5] * u[1] + gc[7] * u[2] + gc[8] * u[3];
  l[3] = gc[6] * u[1] + gc[8] * u[2] + gc[9] * u[3];
}

/* acc^i = Gamma^i_jk K^j K^k (the contraction push_photon forms,
 * src/geodesics.c:22-26) from dg/dX1, dg/dX2. */
IPG_HD void geodesic_accel(const Metric &m, const double K[4], double acc[4])
{
  double v1[4], v2[4];
  lower_g(m.d1, K, v1);
  lower_g(m.d2, K, v2);
  double al[4];
  for (int l = 0; l < 4; l++) al[l] = K[1] * v1[l] + K[2] * v2[l];
  al[1] -= 0.5 * dot4(v1, K);
  al[2] -= 0.5 * dot4(v2, K);
  raise_gc(m.gc, al, acc);
}

/* M^i_k = Gamma^i_kl K^l, the matrix parallel transport of N applies
 * (src/ipolarray.c:281-287). */
IPG_HD void transport_matrix(const Metric &m, const double K[4], double M[4][4])
{
  double v1[4], v2[4];
  lower_g(m.d1, K, v1);
  lower_g(m.d2, K, v2);
  double Ml[4][4];
  for (int l = 0; l < 4; l++)
    for (int k = 0; k < 4; k++) {
      /* dg_02, dg_23 vanish (lower_g) */
      const bool z = sy(l, k) == 2 || sy(l, k) == 8;
      double t = z ? 0. : K[1] * m.d1[sy(l, k)] + K[2] * m.d2[sy(l, k)];
      if (k == 1) t += v1[l];
      if (k == 2) t += v2[l];
      if (l == 1) t -= v1[k];
      if (l == 2) t -= v2[k];
      Ml[l][k] = 0.5 * t;
    }
  for (int k = 0; k < 4; k++) {
    const double col[4] = {Ml[0][k], Ml[1][k], Ml[2][k], Ml[3][k]};
    double up[4];
    raise_gc(m.gc, col, up);
    for (int i = 0; i < 4; i++) M[i][k] = up[i];
  }
}

/* ------------------------------------------------------------- geodesics */

/* src/model_geodesics.c:267-293 */
IPG_HD double stepsize(const ipgpu_geom &G, const double X[4], const double K[4], double eps)
{
  /* 1 / (|dlx| + SS) with dlx = eps f / (|K| + SS) >= 0 is written
   * (|K| + SS) / (eps f + SS (|K| + SS)): one reciprocal per direction
   * instead of two (the same value up to roundoff) */
  const double SS = SMALL * SMALL;
  double deh = fmin(fabs(X[1] - G.cstartx[1]), 0.1);
  double cut = 0.02;
  double ilx2 = rcp(G.cstopx[2] - G.cstartx[2]);
  double dpole = fmin(fabs(X[2] * ilx2), fabs((G.cstopx[2] - X[2]) * ilx2));
  double d2fac = (dpole < cut) ? dpole * (1. / 3.) : fmin(cut / 3 + (dpole - cut) * 10., 1);
  const double a1 = fabs(K[1]) + SS, a2 = fabs(K[2]) + SS, a3 = fabs(K[3]) + SS;
  double idlx1 = a1 * rcp(eps * (10 * deh) + SS * a1);
  double idlx2 = a2 * rcp(eps * d2fac + SS * a2);
  double idlx3 = a3 * rcp(eps + SS * a3);
  return rcp(idlx1 + idlx2 + idlx3);
}

/* src/model_geodesics.c:218-237 */
IPG_HD bool stop_backward(const ipgpu_geom &G, double r, const double K[4])
{
  return (r > G.rmax_geo && K[1] < 0.) || r < (G.Rh + 0.0001);
}

/* First half of push_photon (src/geodesics.c:15-31): from (X, K) at the
 * start of a step of length dlp (negative when tracing backward). */
IPG_HD_INLINE void half_step_acc(const double acc[4], const double X[4], const double K[4], double dlp,
                                  double Xh[4], double Kh[4])
{
  for (int k = 0; k < 4; k++) Kh[k] = K[k] - 0.5 * dlp * acc[k];
  for (int i = 0; i < 4; i++) Xh[i] = X[i] + 0.5 * dlp * K[i];
}

IPG_HD void half_step(const Metric &m0, const double X[4], const double K[4], double dlp,
                      double Xh[4], double Kh[4])
{
  double acc[4];
  geodesic_accel(m0, K, acc);
  half_step_acc(acc, X, K, dlp, Xh, Kh);
}

/* --------------------------------------------------------------- tetrads */

IPG_HD double gdot(const double g[10], const double a[4], const double b[4])
{
  double bl[4];
  lower(g, b, bl);
  return dot4(a, bl);
}

IPG_HD void normalize(double v[4], const double g[10])
{
  double in = rcp(sqrt(fabs(gdot(g, v, v))));
  for (int k = 0; k < 4; k++) v[k] *= in;
}

/* src/tetrads.c:95-115 */
IPG_HD void project_out(double a[4], const double b[4], const double g[10])
{
  double bb = gdot(g, b, b);
  double ab = gdot(g, a, b);
  double f = ab * rcp(bb);
  for (int k = 0; k < 4; k++) a[k] -= b[k] * f;
}

IPG_HD void set_from_trial(double E[4], int defdir, const double t[4])
{
  double n = fabs(t[1]) + fabs(t[2]) + fabs(t[3]);
  for (int k = 0; k < 4; k++) E[k] = (n <= SMALL) ? (double)(k == defdir) : t[k];
}

IPG_HD double det4(const double E[4][4])
{
  double s0 = E[0][0] * E[1][1] - E[1][0] * E[0][1];
  double s1 = E[0][0] * E[1][2] - E[1][0] * E[0][2];
  double s2 = E[0][0] * E[1][3] - E[1][0] * E[0][3];
  double s3 = E[0][1] * E[1][2] - E[1][1] * E[0][2];
  double s4 = E[0][1] * E[1][3] - E[1][1] * E[0][3];
  double s5 = E[0][2] * E[1][3] - E[1][2] * E[0][3];
  double c5 = E[2][2] * E[3][3] - E[3][2] * E[2][3];
  double c4 = E[2][1] * E[3][3] - E[3][1] * E[2][3];
  double c3 = E[2][1] * E[3][2] - E[3][1] * E[2][2];
  double c2 = E[2][0] * E[3][3] - E[3][0] * E[2][3];
  double c1 = E[2][0] * E[3][2] - E[3][0] * E[2][2];
  double c0 = E[2][0] * E[3][1] - E[3][0] * E[2][1];
  return s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
}

/* Plasma-frame tetrad, src/model_tetrads.c:35-127: e0 || U, e3 || K,
 * e2 || B, e1 completes a right-handed set; the same Gram-Schmidt order with
 * one re-orthogonalization pass.  Returns the oddflag bits (1: handedness
 * product off unit by more than 1e-10).  E is Econ[a][mu]. */
IPG_HD int plasma_tetrad(const double U[4], const double K[4], const double B[4], const double g[10],
                         double sqrtmg, double E[4][4])
{
  const double ones[4] = {1., 1., 1., 1.};
  set_from_trial(E[0], 0, U);
  set_from_trial(E[1], 3, ones);
  set_from_trial(E[2], 2, B);
  set_from_trial(E[3], 3, K);
  normalize(E[0], g);
  project_out(E[3], E[0], g);
  project_out(E[3], E[0], g);
  normalize(E[3], g);
  project_out(E[2], E[0], g);
  project_out(E[2], E[3], g);
  project_out(E[2], E[0], g);
  project_out(E[2], E[3], g);
  normalize(E[2], g);
  project_out(E[1], E[0], g);
  project_out(E[1], E[2], g);
  project_out(E[1], E[3], g);
  project_out(E[1], E[0], g);
  project_out(E[1], E[2], g);
  project_out(E[1], E[3], g);
  normalize(E[1], g);
  double dot = sqrtmg * det4(E);
  int flag = 0;
  if (fabs(fabs(dot) - 1.) > 1.e-10) flag |= 1;
  if (dot < 0.)
    for (int k = 0; k < 4; k++) E[1][k] *= -1.;
  return flag;
}

/* Camera tetrad at the camera position, src/model_tetrads.c:143-193 and the
 * old centering of :209-249. */
IPG_HD int camera_tetrad(const ipgpu_geom &G, const double Xcam[4], int old_centering,
                         double E[4][4], Metric &m)
{
  metric(G, Xcam[1], Xcam[2], m);
  double U[4], K[4];
  if (old_centering) {
    U[0] = 1.;
    U[1] = 0.;
    U[2] = 0.;
    U[3] = 0.;
  } else {
    for (int i = 0; i < 4; i++) U[i] = -m.gc[sy(i, 0)];
  }
  for (int i = 0; i < 4; i++) K[i] = m.gc[sy(i, 0)] + m.gc[sy(i, 1)];
  const double trial[4] = {0., 0., 1., 0.};
  return plasma_tetrad(U, K, trial, m.g, m.sqrtmg, E);
}

/* Wavevector of pixel (i, j), src/model_geodesics.c:178-212 with
 * null_normalize (src/geometry.c:190-201) and the K *= freq of
 * src/main.c:837. */
IPG_HD void init_K(const ipgpu_image &P, const double E[4][4], long i, long j, double K[4])
{
  double dxoff = (i + 0.5 + P.xoff - 0.01) / P.nx - 0.5;
  double dyoff = (j + 0.5 + P.yoff) / P.ny - 0.5;
  double Kt[4];
  Kt[0] = 0.;
  Kt[1] = (dxoff * cos(P.rotcam) - dyoff * sin(P.rotcam)) * P.fovx;
  Kt[2] = (dxoff * sin(P.rotcam) + dyoff * cos(P.rotcam)) * P.fovy;
  Kt[3] = 1.;
  double inorm = sqrt(Kt[1] * Kt[1] + Kt[2] * Kt[2] + Kt[3] * Kt[3]);
  Kt[0] = 1.;
  Kt[1] *= 1. / inorm;
  Kt[2] *= 1. / inorm;
  Kt[3] *= 1. / inorm;
  for (int l = 0; l < 4; l++)
    K[l] = (E[0][l] * Kt[0] + E[1][l] * Kt[1] + E[2][l] * Kt[2] + E[3][l] * Kt[3]) * P.freq;
}

/* ------------------------------------------------------------ fluid model */

struct Fluid {
  double Ucon[4], Ucov[4], Bcon[4], Bcov[4];
  double ne, thetae, B;
};

/* Ghosted-grid trilinear weights, src/grid.c:216-283 (Xtoijk_ghost) */
struct Cell {
  int i, j, k;
  double d1, d2, d3;
};

IPG_HD void locate(const ipgpu_geom &G, double X1, double X2, double X3, Cell &c)
{
  double phi = fmod(X3, G.cstopx[3]);
  if (phi < 0.0) phi = G.cstopx[3] + phi;
  const double idx1 = rcp(G.dx[1]), idx2 = rcp(G.dx[2]), idx3 = rcp(G.dx[3]);
  double fi = (X1 - G.startx[1]) * idx1 - 0.5 + 1000;
  double fj = (X2 - G.startx[2]) * idx2 - 0.5 + 1000;
  double fk = (phi - G.startx[3]) * idx3 - 0.5 + 1000;
  /* clamp before the conversion so out-of-range or NaN positions stay in
   * bounds; inside the range this is the reference's (int) truncation */
  fi = fmin(fmax(fi, 0.), 1.e9);
  fj = fmin(fmax(fj, 0.), 1.e9);
  fk = fmin(fmax(fk, 0.), 1.e9);
  int i = (int)fi - 1000, j = (int)fj - 1000, k = (int)fk - 1000;
  if (i < -1) i = -1;
  if (j < -1) j = -1;
  if (k < -1) k = -1;
  if (i >= G.N1) i = G.N1 - 1;
  if (j >= G.N2) j = G.N2 - 1;
  if (k >= G.N3) k = G.N3 - 1;
  double d1 = (X1 - ((i + 0.5) * G.dx[1] + G.startx[1])) * idx1;
  double d2 = (X2 - ((j + 0.5) * G.dx[2] + G.startx[2])) * idx2;
  double d3 = (phi - ((k + 0.5) * G.dx[3] + G.startx[3])) * idx3;
  if (d1 < 0.) d1 = 0.;
  if (d1 >= 1.) d1 = 1.;
  if (d2 < 0.) d2 = 0.;
  if (d2 >= 1.) d2 = 1.;
  if (d3 < 0.) d3 = 0.;
  if (d3 >= 1.) d3 = 1.;
  c.i = i + 1;
  c.j = j + 1;
  c.k = k + 1;
  c.d1 = d1;
  c.d2 = d2;
  c.d3 = d3;
}

IPG_HD bool in_domain(const ipgpu_geom &G, double X1, double X2)
{
  return !(X1 < G.startx[1] || X1 >= G.stopx[1] || X2 < G.startx[2] || X2 >= G.stopx[2]);
}

/* Cell record of the device dump: 16 doubles (one 128-byte line) holding
 * U1 U2 U3 B1 B2 B3 ne thetae b sigma and 6 unused slots. */
constexpr int CELL_STRIDE = 16;

/* Trilinear interpolation of the 10 fields at once (src/grid.c:15-44).
 * at(n, f) is field f of corner n = 4 di + 2 dj + dk, the cell record at
 * (i + di, j + dj, k + dk); one body for every source of the corner
 * records, with its multiply-adds written out (the compiler's own fusing
 * of a*b + c*d depends on the code around it), so that every source gives
 * the same roundings. */
template <class At>
IPG_HD_INLINE void interp10(const Cell &c, const At &at, double v[10])
{
  const double b1 = 1. - c.d1, b2 = 1. - c.d2;
  const double w00 = b1 * b2, w01 = b1 * c.d2, w10 = c.d1 * b2, w11 = c.d1 * c.d2;
  const double z0 = 1. - c.d3, z1 = c.d3;
  for (int f = 0; f < 10; f++) {
    const double lo = fma(at(6, f), w11, fma(at(4, f), w10, fma(at(2, f), w01, at(0, f) * w00)));
    const double hi = fma(at(7, f), w11, fma(at(5, f), w10, fma(at(3, f), w01, at(1, f) * w00)));
    v[f] = fma(z1, hi, z0 * lo);
  }
}

/* The corner records straight from the dump. */
IPG_HD void gather10(const ipgpu_geom &G, const double *dump, const Cell &c, double v[10])
{
  const long s2 = (G.N3 + 2), s1 = (long)(G.N2 + 2) * (G.N3 + 2);
  const double *p = dump + CELL_STRIDE * ((long)c.i * s1 + (long)c.j * s2 + c.k);
  const long o[8] = {0, CELL_STRIDE, CELL_STRIDE * s2, CELL_STRIDE * (s2 + 1), CELL_STRIDE * s1,
                     CELL_STRIDE * (s1 + 1), CELL_STRIDE * (s1 + s2), CELL_STRIDE * (s1 + s2 + 1)};
  interp10(c, [&](int n, int f) { return p[o[n] + f]; }, v);
}

/* The cell records' 10 fields rounded to FP32, 16 floats (64 bytes) per
 * cell: what the mixed-precision batch transfer gathers (half the bytes and
 * registers of the FP64 records; the interpolation stays FP64). */
constexpr int CELL32 = 16;

/* gather10 over the FP32 records (the same corners and interpolation). */
IPG_HD void gather10_f32(const ipgpu_geom &G, const float *d32, const Cell &c, double v[10])
{
  const long s2 = (G.N3 + 2), s1 = (long)(G.N2 + 2) * (G.N3 + 2);
  const float *p = d32 + CELL32 * ((long)c.i * s1 + (long)c.j * s2 + c.k);
  const long o[8] = {0, CELL32, CELL32 * s2, CELL32 * (s2 + 1), CELL32 * s1,
                     CELL32 * (s1 + 1), CELL32 * (s1 + s2), CELL32 * (s1 + s2 + 1)};
  interp10(c, [&](int n, int f) { return (double)p[o[n] + f]; }, v);
}

/* model/iharm/model.c:567-575 */
IPG_HD double sigma_smoothfac(const ipgpu_model &M, double sigma)
{
  double above = M.sigma_cut;
  if (M.sigma_cut_high > 0) above = M.sigma_cut_high;
  if (sigma < M.sigma_cut) return 1;
  if (sigma >= above) return 0;
  double dsig = above - M.sigma_cut;
  return cos(PI / 2. / dsig * (sigma - M.sigma_cut));
}

IPG_HD bool radiating_region(const ipgpu_geom &G, const ipgpu_model &M, double r, double th)
{
  if (M.kind == IPGPU_MODEL_SPHERE) return r < M.sph_R0;       /* model/sphere/model.c:213-218 */
  return (r > G.rmin_geo && r < G.rmax_geo && th > M.th_beg     /* model/iharm/model.c:2093-2098 */
          && th < (PI - M.th_beg));
}

/* One slot of the 8 cell records around a point, trilinear (src/grid.c:15-44). */
IPG_HD double gather_slot(const ipgpu_geom &G, const double *dump, const Cell &c, int f)
{
  const long s2 = (G.N3 + 2), s1 = (long)(G.N2 + 2) * (G.N3 + 2);
  const long base = (long)c.i * s1 + (long)c.j * s2 + c.k;
  const double b1 = 1. - c.d1, b2 = 1. - c.d2;
  const double *p000 = dump + CELL_STRIDE * base;
  const double *p010 = dump + CELL_STRIDE * (base + s2);
  const double *p100 = dump + CELL_STRIDE * (base + s1);
  const double *p110 = dump + CELL_STRIDE * (base + s1 + s2);
  double lo = p000[f] * b1 * b2 + p010[f] * b1 * c.d2 + p100[f] * c.d1 * b2 + p110[f] * c.d1 * c.d2;
  double hi = p000[CELL_STRIDE + f] * b1 * b2 + p010[CELL_STRIDE + f] * b1 * c.d2
      + p100[CELL_STRIDE + f] * c.d1 * b2 + p110[CELL_STRIDE + f] * c.d1 * c.d2;
  return (1. - c.d3) * lo + c.d3 * hi;
}

/* slots of a cell record beyond the 10 fields (filled by the derivation) */
constexpr int SLOT_RHO = 10, SLOT_U = 11, SLOT_BETA = 12, SLOT_DTH_RHIGH = 13, SLOT_DTH_RLOW = 14;

/* What the adjoint needs besides the fluid state: the interpolation cell
 * (ghosted linear index of its lower corner, -1 outside the grid) and
 * weights, the sigma smoothing factor, and dTheta_e/dR_high, dR_low at the
 * point. */
struct FluidExt {
  double cell, d1, d2, d3, smooth, dth_rh, dth_rl;
};

/* Fluid state at a point: get_model_fourv/ne/thetae/b in one evaluation
 * (model/iharm/model.c:420-593; model/sphere/model.c:126-211).  With
 * use_vpre, vpre holds the 10 interpolated fields at X that the caller has
 * formed already (the values gather10 would return). */
IPG_HD void fluid(const ipgpu_geom &G, const ipgpu_model &M, const double *dump, const double X[4],
                  const Metric &m, Fluid &F, FluidExt *ext = nullptr, const double *vpre = nullptr,
                  bool use_vpre = false)
{
  if (ext) {
    ext->cell = -1.;
    ext->d1 = ext->d2 = ext->d3 = ext->smooth = ext->dth_rh = ext->dth_rl = 0.;
  }
  if (M.kind == IPGPU_MODEL_SPHERE) {
    double U[4] = {1., 0., 0., 0.};
    double B[4] = {0., M.sph_B0 * cos(m.th) / m.r, -M.sph_B0 * sin(m.th) / (m.r + 1.e-8), 0.};
    for (int i = 0; i < 4; i++) {
      F.Ucon[i] = U[i];
      F.Bcon[i] = B[i];
    }
    lower(m.g, F.Ucon, F.Ucov);
    lower(m.g, F.Bcon, F.Bcov);
    bool rad = m.r < M.sph_R0;
    F.ne = rad ? M.sph_Ne0 : 0.;
    F.thetae = rad ? M.sph_Thetae0 : 0.;
    F.B = rad ? M.sph_B0 : 0.;
    return;
  }
  if (!in_domain(G, X[1], X[2])) {
    F.Ucov[0] = -1. / sqrt(-m.gc[0]);
    F.Ucov[1] = 0.;
    F.Ucov[2] = 0.;
    F.Ucov[3] = 0.;
    for (int i = 0; i < 4; i++) F.Ucon[i] = m.gc[sy(i, 0)] * F.Ucov[0];
    for (int i = 0; i < 4; i++) {
      F.Bcon[i] = 0.;
      F.Bcov[i] = 0.;
    }
    F.ne = 0.;
    F.thetae = 0.;
    F.B = 0.;
    return;
  }
  Cell c;
  locate(G, X[1], X[2], X[3], c);
  double v[10];
  if (use_vpre)
    for (int f = 0; f < 10; f++) v[f] = vpre[f];
  else
    gather10(G, dump, c, v);
  double VdotV = 0.;
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) VdotV += m.g[sy(i, j)] * v[i - 1] * v[j - 1];
  double Vfac = sqrt(-rcp(m.gc[0]) * (1. + fabs(VdotV)));
  F.Ucon[0] = -Vfac * m.gc[0];
  for (int i = 1; i < 4; i++) F.Ucon[i] = v[i - 1] - Vfac * m.gc[sy(0, i)];
  lower(m.g, F.Ucon, F.Ucov);
  F.Bcon[0] = v[3] * F.Ucov[1] + v[4] * F.Ucov[2] + v[5] * F.Ucov[3];
  const double iu0 = rcp(F.Ucon[0]);
  F.Bcon[1] = (v[3] + F.Ucon[1] * F.Bcon[0]) * iu0;
  F.Bcon[2] = (v[4] + F.Ucon[2] * F.Bcon[0]) * iu0;
  F.Bcon[3] = (v[5] + F.Ucon[3] * F.Bcon[0]) * iu0;
  lower(m.g, F.Bcon, F.Bcov);
  double sigma = v[9];
  const double smooth = (sigma > M.sigma_cut) ? 0. : sigma_smoothfac(M, sigma);
  F.ne = (sigma > M.sigma_cut) ? 0. : v[6] * smooth;
  F.thetae = v[7];
  F.B = v[8];
  if (ext) {
    ext->cell = (double)(((long)c.i * (G.N2 + 2) + c.j) * (long)(G.N3 + 2) + c.k);
    ext->d1 = c.d1;
    ext->d2 = c.d2;
    ext->d3 = c.d3;
    ext->smooth = smooth;
    ext->dth_rh = gather_slot(G, dump, c, SLOT_DTH_RHIGH);
    ext->dth_rl = gather_slot(G, dump, c, SLOT_DTH_RLOW);
  }
}

/* --------------------------------------------------------- K_n(1/Theta_e) */

/* Exponentially scaled K0, K1, K2 at x: ks[n] = e^x K_n(x), and e^-x.
 * x <= 2: power series (A&S 9.6.10-9.6.13); x > 2: Chebyshev expansion of
 * sqrt(x) e^x K_n(x) in 4/x - 1 (ipg_bessel_coeffs.h); K2 by the upward
 * recurrence K2 = K0 + (2/x) K1 as GSL's K_n. */
/* 1/k^2, 1/(k(k+1)), 1/k for the series terms k = 1..15 (index k) */
struct BesselSeriesTables {
  ipg_f64 ik2[17], ikk1[17], ik[17];
  constexpr BesselSeriesTables() : ik2(), ikk1(), ik()
  {
    for (int k = 1; k < 17; k++) {
      ik2[k] = 1.0 / (ipg_f64)(k * k);
      ikk1[k] = 1.0 / (ipg_f64)(k * (k + 1));
      ik[k] = 1.0 / (ipg_f64)k;
    }
  }
};

IPG_HD void bessel_k012(double x, double ks[3], double &emx)
{
  constexpr BesselSeriesTables T;
  emx = exp_(-x);
  const double ix = rcp(x);
  double k0s, k1s;
  if (x <= 2.0) {
    const double EG = 0.57721566490153286061;
    double y = 0.25 * x * x;
    double t0 = 1.0, t1 = 0.5 * x;
    double I0 = 1.0, I1 = t1, S0 = 0., S1 = t1 * (1.0 - 2.0 * EG);
    double Hk = 0.0, p1 = -EG, p2 = 1.0 - EG;
#pragma unroll
    for (int k = 1; k < 16; k++) {
      t0 *= y * T.ik2[k];
      t1 *= y * T.ikk1[k];
      Hk += T.ik[k];
      p1 += T.ik[k];
      p2 += T.ik[k + 1];
      I0 += t0;
      I1 += t1;
      S0 += t0 * Hk;
      S1 += t1 * (p1 + p2);
    }
    double lg = log_(0.5 * x);
    double ex = rcp(emx);
    k0s = (-(lg + EG) * I0 + S0) * ex;
    k1s = (ix + lg * I1 - 0.5 * S1) * ex;
  } else {
    double t = 4.0 * ix - 1.0, t2 = 2.0 * t;
    double b0 = 0., b1 = 0., b2 = 0.;
    for (int k = IPG_K0_NCHEB - 1; k >= 1; k--) {
      b2 = b1;
      b1 = b0;
      b0 = t2 * b1 - b2 + ipg_k0_cheb[k];
    }
    double f0 = t * b0 - b1 + ipg_k0_cheb[0];
    b0 = b1 = b2 = 0.;
    for (int k = IPG_K1_NCHEB - 1; k >= 1; k--) {
      b2 = b1;
      b1 = b0;
      b0 = t2 * b1 - b2 + ipg_k1_cheb[k];
    }
    double f1 = t * b0 - b1 + ipg_k1_cheb[0];
    double isx = rcp(sqrt(x));
    k0s = f0 * isx;
    k1s = f1 * isx;
  }
  ks[0] = k0s;
  ks[1] = k1s;
  ks[2] = k0s + (2.0 * ix) * k1s;
}

/* ---------------------------------------------------- transfer coefficients */

struct Coef {
  double jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV;
};

/* src/radiation.c:14-23 */
IPG_HD double Bnu_inv(double nu, double Thetae)
{
  double x = (HPL / (ME * CL * CL)) * nu * rcp(Thetae);
  if (x < 2.e-3) return ((2. * HPL / (CL * CL)) * rcp(x * (1. / 24.) * (24. + x * (12. + x * (4. + x)))));
  return ((2. * HPL / (CL * CL)) * rcp(exp_(x) - 1.));
}

/* src/symphony/maxwell_juettner_fits.c:334-356 with x^(1/3) shared */
IPG_HD void dexter_IQV(double x, double &II, double &IQ, double &IV)
{
  double c = cbrt_(x);
  double ic = rcp(c), ic2 = ic * ic;
  double e = exp_(-1.8899 * c);
  II = 2.5651 * (1 + 1.92 * ic + 0.9977 * ic2) * e;
  IQ = 2.5651 * (1 + 0.93193 * ic + 0.499873 * ic2) * e;
  IV = (1.81384 * rcp(x) + 3.42319 * ic2 + 0.0292545 * rcp(sqrt(x)) + 2.03773 * ic) * e;
}

/* constants of the fits, evaluated once */
constexpr double SQRT2 = 1.4142135623730951;
constexpr double SQRT3 = 1.7320508075688772;
constexpr double TWO_11_12 = 1.887748625363387;          /* 2^(11/12) */
constexpr double C_EXTRA = 0.011767707320949835;         /* 2^(-1/3) / 3^(23/6) */
constexpr double LOG120 = 4.787491742782046;
constexpr double COS_28_25 = 0.43568244627671215, SIN_28_25 = 0.900100442176505;
constexpr double COS_PI_2 = 6.123233995736766e-17;       /* cos of the double nearest pi/2 */

/* Thermal (Maxwell-Juettner) coefficients at one point: the pol = 0 branch
 * (Leung j_I, Kirchhoff a_I; src/model_radiation.c:226-237) and, if pol,
 * the pol = 1 branch (Dexter or Pandya j_S, Kirchhoff a_S, Dexter rho_Q and
 * Shcherbakov rho_V; :243-305), sharing the fluid state, the fluid-frame
 * frequency and angle, and one K_n(1/Theta_e) evaluation.  sin and cos of
 * the field angle come from mu = cos(theta) directly (theta = acos(mu) in
 * [0, pi]); the fits' powers share their logarithms.
 * ju/ku: j_I/nu^2 and a_I nu of the unpolarized solve. */
/* Fluid-frame frequency and field angle (get_fluid_nu, get_bk_angle,
 * src/radiation.c:26-72): nu, sin and cos of theta, and whether theta is 0
 * or pi. */
struct Geo {
  double nu, sth, cth;
  bool edge;
};

IPG_HD Geo coef_geometry(const Fluid &F, const double K[4])
{
  Geo q;
  double kdotu = K[0] * F.Ucov[0] + K[1] * F.Ucov[1] + K[2] * F.Ucov[2] + K[3] * F.Ucov[3];
  q.nu = -kdotu * (ME * CL * CL / HPL);
  q.edge = false;
  double Bm = sqrt(fabs(F.Bcon[0] * F.Bcov[0] + F.Bcon[1] * F.Bcov[1] + F.Bcon[2] * F.Bcov[2]
                        + F.Bcon[3] * F.Bcov[3]));
  if (Bm == 0.) {
    q.sth = 1.;
    q.cth = COS_PI_2;
  } else {
    double mu = (K[0] * F.Bcov[0] + K[1] * F.Bcov[1] + K[2] * F.Bcov[2] + K[3] * F.Bcov[3])
        * rcp(fabs(kdotu) * Bm);
    if (fabs(mu) > 1.) mu = (mu > 0.) ? 1. : -1.;
    q.edge = (mu == 1. || mu == -1.);   /* theta = 0 or pi */
    q.sth = sqrt((1. - mu) * (1. + mu));
    q.cth = mu;
  }
  return q;
}

/* The fits proper, from the fluid-frame frequency and angle and the local
 * n_e, Theta_e and B. */
IPG_HD void coefficients_core(const Geo &q, double Ne, double Th, double B, int emission_type, bool pol,
                              double &ju, double &ku, Coef &c)
{
  ju = 0.;
  ku = 0.;
  c.jI = c.jQ = c.jU = c.jV = c.aI = c.aQ = c.aU = c.aV = c.rQ = c.rU = c.rV = 0.;
  if (Ne <= 0.) return;
  const double nu = q.nu, sth = q.sth, cth = q.cth;
  const bool edge = q.edge;
  double inu = rcp(nu);
  double nusq = nu * nu;
  double iTh = rcp(Th);
  double ks[3], emx;
  bessel_k012(iTh, ks, emx);
  /* Faraday fits, src/symphony/maxwell_juettner_fits.c:207-284 */
  double omega0 = (EE / (ME * CL)) * B;
  double wp2 = (4. * PI * EE * EE / ME) * Ne;
  double xr = Th * sqrt(SQRT2 * sth * (1.e3 / (2. * PI)) * omega0 * inu);
  double w = 2. * PI * nu;
  double iw = rcp(w);
  double iks2 = rcp(ks[2]);
  if (edge) {
    if (pol) {
      double fit;
      if (emission_type != 1 && emx * ks[2] > 0) {
        double sh = 0.43793091 * log_(1. + 0.00185777 * pow_(xr, 1.50316886));
        fit = (emx * ks[0] - sh) * rcp(emx * ks[2]);
      } else {
        fit = (ks[0] * iks2) * (1 - 0.11 * log_(1 + 0.035 * xr));
      }
      c.rV = 2. * PI * nu / CL * (wp2 * omega0 * (iw * iw * iw) * fit * cth) * nu;
    }
    return;
  }
  double Bnuinv = Bnu_inv(nu, Th);
  double iBnu = (Bnuinv > 0) ? rcp(Bnuinv) : 0.;
  /* unpolarized: Leung fit, maxwell_juettner_fits.c:368-395 */
  {
    double K2 = fmax(emx * ks[2], SMALL);
    double nuc = (EE / (2. * PI * ME * CL)) * B;
    double nus = (2. / 9.) * nuc * Th * Th * sth;
    double j = 0.;
    if (!(nu > 1.e12 * nus)) {
      double x = nu * rcp(nus);
      double c3 = cbrt_(x);
      double f = sqrt(x) + TWO_11_12 * sqrt(c3);   /* x^(1/6) = sqrt(x^(1/3)) */
      f = f * f;
      j = (SQRT2 * PI * EE * EE / (3. * CL)) * Ne * nus * rcp(K2) * f * exp_(-c3);
    }
    ju = j * rcp(nusq);
    ku = ju * iBnu;
  }
  if (!pol) return;
  double jI, jQ, jV;
  if (emission_type == 1) {
    /* Pandya fits, maxwell_juettner_fits.c:27-129 */
    double nu_c = (EE / (2. * PI * ME * CL)) * B;
    double nu_s = (2. / 9.) * nu_c * sth * Th * Th;
    double X = nu * rcp(nu_s);
    double pre = (EE * EE / CL) * Ne * nu_c;
    double LX = log_(X);
    double X13 = exp_(LX * (1. / 3.)), X16 = sqrt(X13), e13 = exp_(-X13);
    double t1 = SQRT2 * PI / 27. * sth;
    double sX = sqrt(X);
    double tI = sX + TWO_11_12 * X16;
    jI = pre * t1 * (tI * tI) * e13;
    double LT = log_(Th);
    double Tq = exp_(LT * (24. / 25.));
    double tq2 = (7. * Tq + 35.) * rcp(10. * Tq + 75.);
    double tq3 = sX + tq2 * TWO_11_12 * X16;
    jQ = pre * t1 * (tq3 * tq3) * e13;  /* = -maxwell_juettner_Q */
    double sin_t = sth * COS_28_25 - cth * SIN_28_25;   /* sin(theta - 28/25) */
    double tv1 = (37. - 87. * sin_t) * rcp(100. * (Th + 1.));
    double tv2 = pow_(1. + (exp_(LT * (3. / 5.)) * (1. / 25.) + 7. / 10.) * exp_(LX * (9. / 25.)), 5. / 3.);
    jV = pre * tv1 * tv2 * e13;
  } else {
    /* Dexter fits, maxwell_juettner_fits.c:287-330 */
    double nus = (3.0 * EE / (4.0 * PI * ME * CL)) * B * sth * Th * Th + 1.0;
    double x = nu * rcp(nus);
    double II, IQ, IV;
    dexter_IQV(x, II, IQ, IV);
    double iTh2 = iTh * iTh;
    double pre = (EE * EE / (2. * SQRT3 * CL)) * Ne * nu * iTh2;
    jI = pre * II;
    jQ = pre * IQ;
    jV = (2. * EE * EE / (3. * SQRT3 * CL)) * Ne * nu * (cth * rcp(sth)) * iTh2 * iTh * IV;
  }
  /* src/model_radiation.c:245-281 */
  double inu2 = rcp(nusq);
  c.jI = jI * inu2;
  c.jQ = jQ * inu2;
  c.jU = -0. * inu2;
  c.jV = jV * inu2;
  double jP = sqrt(c.jQ * c.jQ + c.jU * c.jU + c.jV * c.jV);
  if (c.jI < jP * (1. / 0.99)) {
    double pf = c.jI * rcp(jP) * 0.99;
    c.jQ *= pf;
    c.jU *= pf;
    c.jV *= pf;
  }
  c.aI = c.jI * iBnu;
  c.aQ = c.jQ * iBnu;
  c.aU = c.jU * iBnu;
  c.aV = c.jV * iBnu;
  /* rho_Q (Dexter 2016) and rho_V (Shcherbakov), dexter_fit = 0 */
  {
    double xe = xr + 1.e-16;
    double Lr = log_(xr);
    double Le = (xe == xr) ? Lr : log_(xe);
    double e472 = exp_(-xr * (1. / 47.2));
    /* 0.5 + 0.5 tanh(z) = 1 / (1 + e^(-2z)) */
    double mz2 = -2. * (Lr - LOG120) * (1. / 0.1);
    double logistic = (mz2 > 700.) ? 0. : rcp(1. + exp_(mz2));
    double extraterm = (.011 * e472 - C_EXTRA * PI * 1.e4 * exp_(Le * (-8. / 3.))) * logistic;
    double sh2, ch2;
    sincos_(0.5 * xr, sh2, ch2);
    double jff = 2.011 * exp_(-exp_(1.035 * Lr) * (1. / 4.7)) - ch2 * exp_(-exp_(1.2 * Lr) * (1. / 2.73))
        - .011 * e472 + extraterm;
    double kr = ks[1] * iks2;
    double iw2 = iw * iw;
    double eps11m22 = jff * wp2 * (omega0 * omega0) * (iw2 * iw2) * (kr + 6. * Th) * (sth * sth);
    c.rQ = (PI / CL) * nu * eps11m22 * nu;
    double fit = (ks[0] * iks2) * (1 - 0.11 * log_(1 + 0.035 * xr));
    double eps12 = wp2 * omega0 * (iw2 * iw) * fit * cth;
    c.rV = (2. * PI / CL) * nu * eps12 * nu;
    c.rU = 0. * nu;
  }
}

/* Coefficients at a point from the fluid state and the wavevector. */
IPG_HD void coefficients(const Fluid &F, const double K[4], int emission_type, bool pol, double &ju,
                         double &ku, Coef &c)
{
  if (F.ne <= 0.) {
    coefficients_core(Geo{0., 1., 0., false}, 0., 1., 0., emission_type, pol, ju, ku, c);
    return;
  }
  coefficients_core(coef_geometry(F, K), F.ne, F.thetae, F.B, emission_type, pol, ju, ku, c);
}

/* ------------------------------------------- coefficients, FP32 shape functions */

/* The mixed-precision variant of coefficients(): the same formulas, with
 * the dimensionless shape functions of the fits (K_n(1/Theta_e), the Dexter
 * and Leung spectra, the Faraday fits and Bnu's exponential) evaluated in
 * FP32, and the fluid-frame frequency, the field angle and the dimensional
 * prefactors (whose magnitudes, ~1e-35 for j/nu^2, lie outside FP32's
 * range) in FP64. */
IPG_HD void bessel_k012_f(float x, float ks[3], float &emx)
{
  emx = expf(-x);
  float k0s, k1s;
  if (x <= 2.0f) {
    const float EG = 0.5772156649f;
    float y = 0.25f * x * x;
    float t0 = 1.0f, t1 = 0.5f * x;
    float I0 = 1.0f, I1 = t1, S0 = 0.f, S1 = t1 * (1.0f - 2.0f * EG);
    float Hk = 0.0f, p1 = -EG, p2 = 1.0f - EG;
#pragma unroll
    for (int k = 1; k < 10; k++) {
      const float ik = 1.0f / (float)k, ik1 = 1.0f / (float)(k + 1);
      t0 *= y * (ik * ik);
      t1 *= y * (ik * ik1);
      Hk += ik;
      p1 += ik;
      p2 += ik1;
      I0 += t0;
      I1 += t1;
      S0 += t0 * Hk;
      S1 += t1 * (p1 + p2);
    }
    float lg = logf(0.5f * x);
    float ex = expf(x);
    k0s = (-(lg + EG) * I0 + S0) * ex;
    k1s = (1.0f / x + lg * I1 - 0.5f * S1) * ex;
  } else {
    float t = 4.0f / x - 1.0f, t2 = 2.0f * t;
    float b0 = 0.f, b1 = 0.f, b2 = 0.f;
    for (int k = 11; k >= 1; k--) {
      b2 = b1;
      b1 = b0;
      b0 = t2 * b1 - b2 + (float)ipg_k0_cheb[k];
    }
    float f0 = t * b0 - b1 + (float)ipg_k0_cheb[0];
    b0 = b1 = b2 = 0.f;
    for (int k = 11; k >= 1; k--) {
      b2 = b1;
      b1 = b0;
      b0 = t2 * b1 - b2 + (float)ipg_k1_cheb[k];
    }
    float f1 = t * b0 - b1 + (float)ipg_k1_cheb[0];
    float isx = 1.0f / sqrtf(x);
    k0s = f0 * isx;
    k1s = f1 * isx;
  }
  ks[0] = k0s;
  ks[1] = k1s;
  ks[2] = k0s + (2.0f / x) * k1s;
}

/* The fits of coefficients_f32 from the fluid-frame frequency and angle
 * (coef_geometry) and the local n_e, Theta_e and B. */
IPG_HD void coefficients_f32_core(const Geo &q, double Ne, double Th, double B, int emission_type, bool pol,
                                  double &ju, double &ku, Coef &c)
{
  ju = 0.;
  ku = 0.;
  c.jI = c.jQ = c.jU = c.jV = c.aI = c.aQ = c.aU = c.aV = c.rQ = c.rU = c.rV = 0.;
  if (Ne <= 0.) return;
  const double nu = q.nu;
  const double nusq = nu * nu;
  const double sth = q.sth, cth = q.cth;
  const bool edge = q.edge;
  const float Thf = (float)Th, sthf = (float)sth, cthf = (float)cth;
  float ks[3], emx;
  bessel_k012_f(1.0f / Thf, ks, emx);
  double omega0 = (EE / (ME * CL)) * B;
  double wp2 = (4. * PI * EE * EE / ME) * Ne;
  const float xr = (float)(Th * sqrt(SQRT2 * sth * (1.e3 / (2. * PI)) * omega0 / nu));
  double w = 2. * PI * nu;
  if (edge) {
    if (pol) {
      float fit;
      if (emission_type != 1 && emx * ks[2] > 0) {
        float sh = 0.43793091f * logf(1.f + 0.00185777f * exp2f(1.50316886f * log2f(xr)));
        fit = (emx * ks[0] - sh) / (emx * ks[2]);
      } else {
        fit = (ks[0] / ks[2]) * (1.f - 0.11f * logf(1.f + 0.035f * xr));
      }
      c.rV = 2. * PI * nu / CL * (wp2 * omega0 / (w * w * w) * (double)fit * cth) * nu;
    }
    return;
  }
  /* Bnu_inv (src/radiation.c:14-23): the prefactor in FP64, e^x - 1 in FP32 */
  double Bnuinv;
  {
    const float x = (float)((HPL / (ME * CL * CL)) * nu / Th);
    if (x < 2.e-3f) Bnuinv = (2. * HPL / (CL * CL)) / (double)(x / 24.f * (24.f + x * (12.f + x * (4.f + x))));
    else Bnuinv = (2. * HPL / (CL * CL)) / (double)expm1f(x);
  }
  double iBnu = (Bnuinv > 0) ? 1. / Bnuinv : 0.;
  {
    double K2 = fmax((double)emx * (double)ks[2], SMALL);
    double nuc = (EE / (2. * PI * ME * CL)) * B;
    double nus = (2. / 9.) * nuc * Th * Th * sth;
    double j = 0.;
    if (!(nu > 1.e12 * nus)) {
      const float x = (float)(nu / nus);
      const float c3 = cbrtf(x);
      float f = sqrtf(x) + 1.887748625f * sqrtf(c3);
      f = f * f;
      j = (SQRT2 * PI * EE * EE / (3. * CL)) * Ne * nus / K2 * (double)(f * expf(-c3));
    }
    ju = j / nusq;
    ku = ju * iBnu;
  }
  if (!pol) return;
  double jI, jQ, jV;
  if (emission_type == 1) {
    double nu_c = (EE / (2. * PI * ME * CL)) * B;
    double nu_s = (2. / 9.) * nu_c * sth * Th * Th;
    const float X = (float)(nu / nu_s);
    double pre = (EE * EE / CL) * Ne * nu_c;
    const float LX = log2f(X), LT = log2f(Thf);
    const float X13 = exp2f(LX * (1.f / 3.f)), X16 = sqrtf(X13), e13 = expf(-X13);
    const float t1 = 1.414213562f * 3.141592654f / 27.f * sthf;
    const float sX = sqrtf(X);
    const float tI = sX + 1.887748625f * X16;
    jI = pre * (double)(t1 * (tI * tI) * e13);
    const float Tq = exp2f(LT * (24.f / 25.f));
    const float tq2 = (7.f * Tq + 35.f) / (10.f * Tq + 75.f);
    const float tq3 = sX + tq2 * 1.887748625f * X16;
    jQ = pre * (double)(t1 * (tq3 * tq3) * e13);
    const float sin_t = sthf * 0.4356824463f - cthf * 0.9001004422f;
    const float tv1 = (37.f - 87.f * sin_t) / (100.f * (Thf + 1.f));
    const float tv2 = exp2f((5.f / 3.f) * log2f(1.f + (exp2f(LT * 0.6f) * 0.04f + 0.7f) * exp2f(LX * 0.36f)));
    jV = pre * (double)(tv1 * tv2 * e13);
  } else {
    double nus = (3.0 * EE / (4.0 * PI * ME * CL)) * B * sth * Th * Th + 1.0;
    const float x = (float)(nu / nus);
    const float cb = cbrtf(x), ic = 1.f / cb, ic2 = ic * ic;
    const float e = expf(-1.8899f * cb);
    const float II = 2.5651f * (1.f + 1.92f * ic + 0.9977f * ic2) * e;
    const float IQ = 2.5651f * (1.f + 0.93193f * ic + 0.499873f * ic2) * e;
    const float IV = (1.81384f / x + 3.42319f * ic2 + 0.0292545f / sqrtf(x) + 2.03773f * ic) * e;
    double iTh2 = 1. / (Th * Th);
    double pre = (EE * EE / (2. * SQRT3 * CL)) * Ne * nu * iTh2;
    jI = pre * (double)II;
    jQ = pre * (double)IQ;
    jV = (2. * EE * EE / (3. * SQRT3 * CL)) * Ne * nu * (cth / sth) * iTh2 / Th * (double)IV;
  }
  double inu2 = 1. / nusq;
  c.jI = jI * inu2;
  c.jQ = jQ * inu2;
  c.jU = -0. * inu2;
  c.jV = jV * inu2;
  double jP = sqrt(c.jQ * c.jQ + c.jU * c.jU + c.jV * c.jV);
  if (c.jI < jP * (1. / 0.99)) {
    double pf = c.jI / jP * 0.99;
    c.jQ *= pf;
    c.jU *= pf;
    c.jV *= pf;
  }
  c.aI = c.jI * iBnu;
  c.aQ = c.jQ * iBnu;
  c.aU = c.jU * iBnu;
  c.aV = c.jV * iBnu;
  {
    const float Lr = logf(xr);
    const float Le = logf(xr + 1.e-16f);
    const float e472 = expf(-xr * (1.f / 47.2f));
    const float mz2 = -2.f * (Lr - 4.787491743f) * 10.f;
    const float logistic = (mz2 > 80.f) ? 0.f : 1.f / (1.f + expf(mz2));
    const float extraterm = (.011f * e472 - 0.01176770732f * 3.141592654f * 1.e4f * expf(Le * (-8.f / 3.f))) * logistic;
    const float jff = 2.011f * expf(-expf(1.035f * Lr) * (1.f / 4.7f)) - cosf(0.5f * xr) * expf(-expf(1.2f * Lr) * (1.f / 2.73f))
        - .011f * e472 + extraterm;
    const double kr = (double)(ks[1] / ks[2]);
    double w2 = w * w;
    double eps11m22 = (double)jff * wp2 * (omega0 * omega0) / (w2 * w2) * (kr + 6. * Th) * (sth * sth);
    c.rQ = (PI / CL) * nu * eps11m22 * nu;
    const double fit = (double)((ks[0] / ks[2]) * (1.f - 0.11f * logf(1.f + 0.035f * xr)));
    double eps12 = wp2 * omega0 / (w2 * w) * fit * cth;
    c.rV = (2. * PI / CL) * nu * eps12 * nu;
    c.rU = 0. * nu;
  }
}

IPG_HD void coefficients_f32(const Fluid &F, const double K[4], int emission_type, bool pol, double &ju,
                             double &ku, Coef &c)
{
  if (F.ne <= 0.) {
    coefficients_f32_core(Geo{0., 1., 0., false}, 0., 1., 0., emission_type, pol, ju, ku, c);
    return;
  }
  coefficients_f32_core(coef_geometry(F, K), F.ne, F.thetae, F.B, emission_type, pol, ju, ku, c);
}

/* ------------------------------------------------------------ radiative step */

/* src/ipolarray.c:607-628 */
IPG_HD double approximate_solve(double Ii, double ji, double ki, double jf, double kf, double dl,
                                double &tau)
{
  double javg = (ji + jf) / 2.;
  double kavg = (ki + kf) / 2.;
  double dtau = dl * kavg;
  tau += dtau;
  if (dtau < 1.e-3) return Ii + (javg - Ii * kavg) * dl * (1. - (dtau * 0.5) * (1. - dtau * (1. / 3.)));
  double efac = exp_neg(-dtau);
  return Ii * efac + (javg * rcp(kavg)) * (1. - efac);
}

/* The plasma-frame Stokes step of evolve_N (Faraday half step, constant-
 * coefficient absorption/emission, Faraday half step; src/ipolarray.c:
 * 391-520) as the affine map S -> T S + cv it is.  T and cv depend only on
 * the coefficients and x, so they are formed wherever the coefficients are
 * (once per step, off the state recurrence); composing the three factors
 * reorders the reference's roundoff only.  The branches on rho^2, aP^2
 * (1e-80), the small-optical-depth series (1e-5) and approximate_solve's
 * 1e-3 switch are the reference's. */
IPG_HD void faraday_matrix(const Coef &c, double x, double R[3][3])
{
  const double r[3] = {c.rQ, c.rU, c.rV};
  const double rho2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  /* [r]x S = r cross S on (Q, U, V) */
  const double X[3][3] = {{0., -r[2], r[1]}, {r[2], 0., -r[0]}, {-r[1], r[0], 0.}};
  if (rho2 > 1e-80) {
    double rho = sqrt(rho2);
    double irho = rcp(rho);
    /* half angle: sh = sin(rho x / 2), then sin and cos of rho x by the
     * double-angle identities */
    double sh, ch;
    sincos_(0.5 * rho * x, sh, ch);
    double s = 2. * sh * ch, cc = 1. - 2. * sh * sh;
    double w = 2. * sh * sh * (irho * irho), v = s * irho;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) R[i][j] = (i == j ? cc : 0.) + w * r[i] * r[j] + v * X[i][j];
  } else {
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) R[i][j] = (i == j ? 1. : 0.) + x * X[i][j];
  }
}

IPG_HD void absorb_emit_affine(const Coef &c, double x, double A[4][4], double b[4])
{
  const double j[4] = {c.jI, c.jQ, c.jU, c.jV};
  const double a[3] = {c.aQ, c.aU, c.aV};
  const double aI = c.aI;
  const double aP2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) A[p][q] = 0.;
  if (aP2 > 1e-80) {
    double aP = sqrt(aP2);
    double iaP = rcp(aP), iaI = rcp(aI);
    double tauP = aP * x, tauI = aI * x;
    double efacm = exp_(-tauI + tauP), efacp = exp_(-tauI - tauP), efac = exp_(-tauI);
    double ch = 0.5 * (efacm + efacp), shp = (efacp - efacm) * (0.5 * iaP);
    double iaP2 = iaP * iaP;
    A[0][0] = ch;
    for (int k = 0; k < 3; k++) {
      A[0][k + 1] = shp * a[k];
      A[k + 1][0] = shp * a[k];
      for (int l = 0; l < 3; l++) A[k + 1][l + 1] = (k == l ? efac : 0.) + a[k] * a[l] * (ch - efac) * iaP2;
    }
    double afacm = 1. - efacm, afac = 1. - efac, afacp = 1. - efacp;
    if (tauI - tauP <= 1e-5) {
      double e = tauI - tauP;
      afacm = e * (1. - (e / 2.) * (1. - e / 3.));
      if (tauI <= 1e-5) {
        e = tauI;
        afac = e * (1. - (e / 2.) * (1. - e / 3.));
        if (tauI + tauP <= 1e-5) {
          e = tauI + tauP;
          afacp = e * (1. - (e / 2.) * (1. - e / 3.));
        }
      }
    }
    double adj = a[0] * j[1] + a[1] * j[2] + a[2] * j[3];
    double m = afacm * (aP * j[0] - adj) * (0.5 * iaP) * rcp(aI - aP);
    double p = afacp * (aP * j[0] + adj) * (0.5 * iaP) * rcp(aI + aP);
    b[0] = m + p;
    for (int k = 0; k < 3; k++)
      b[k + 1] = afac * (j[k + 1] * iaI - a[k] * adj * (iaI * iaP2)) + (p - m) * a[k] * iaP;
  } else {
    double dtau = x * aI;
    if (dtau < 1.e-3) {
      double g = x * (1. - (dtau * 0.5) * (1. - dtau * (1. / 3.)));
      for (int k = 0; k < 4; k++) {
        A[k][k] = 1. - aI * g;
        b[k] = j[k] * g;
      }
    } else {
      double efac = exp_(-dtau);
      double iaI = rcp(aI);
      for (int k = 0; k < 4; k++) {
        A[k][k] = efac;
        b[k] = (j[k] * iaI) * (1. - efac);
      }
    }
  }
}

IPG_HD void stokes_affine(const Coef &c, double x, double T[4][4], double cv[4])
{
  double R3[3][3], A[4][4], b[4];
  faraday_matrix(c, 0.5 * x, R3);
  absorb_emit_affine(c, x, A, b);
  double R[4][4];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 4; k++) R[i][k] = (i == 0 || k == 0) ? (i == k ? 1. : 0.) : R3[i - 1][k - 1];
  double RA[4][4];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 4; k++) {
      double s = 0.;
      for (int l = 0; l < 4; l++) s += R[i][l] * A[l][k];
      RA[i][k] = s;
    }
  for (int i = 0; i < 4; i++) {
    for (int k = 0; k < 4; k++) {
      double s = 0.;
      for (int l = 0; l < 4; l++) s += RA[i][l] * R[l][k];
      T[i][k] = s;
    }
    double s = 0.;
    for (int l = 0; l < 4; l++) s += R[i][l] * b[l];
    cv[i] = s;
  }
}

/* (1 - e^-d) / d from em = e^-d - 1, with the thin-step series of
 * approximate_solve below d = 1e-3 (src/ipolarray.c:616-620) */
IPG_HD float one_minus_exp_over(float d, float em)
{
  return d < 1.e-3f ? 1.f - (d * 0.5f) * (1.f - d * (1.f / 3.f)) : -em / d;
}

/* stokes_affine in FP32: the Faraday rotation, the absorption/emission
 * solution and their products (T and cv) are formed in single precision
 * from dimensionless inputs.  The coefficients are turned into optical
 * depths and rotation angles of the step (a x, rho x: formed in FP64, then
 * rounded) and unit vectors, the emission into j x scaled by a power of two
 * 2^-e that brings its largest component near 1 (j / nu^2 x spans
 * ~1e-40..1e-20 in these units, outside FP32's normal range); sums and
 * differences of optical depths (tau_I -+ tau_P) are taken in FP64 before
 * rounding, and the source terms are written with (1 - e^-tau) / tau so that
 * no product of two small depths (which would underflow) is formed.  cv is
 * scaled back by 2^e exactly.  The branch tests (rho^2, aP^2 against 1e-80; the thin-step
 * switch at 1e-3) are the reference's, on the FP64 coefficients. */
IPG_HD void stokes_affine_f32(const Coef &c, double x, double T[4][4], double cv[4])
{
  /* Faraday half step: angle vector rho x / 2 */
  float R3[3][3];
  {
    const double rho2 = c.rQ * c.rQ + c.rU * c.rU + c.rV * c.rV;
    const double hx = 0.5 * x;
    const float t[3] = {(float)(c.rQ * hx), (float)(c.rU * hx), (float)(c.rV * hx)};
    const float X[3][3] = {{0.f, -t[2], t[1]}, {t[2], 0.f, -t[0]}, {-t[1], t[0], 0.f}};
    const float th = (float)(sqrt(rho2) * hx);     /* rotation angle of the half step */
    if (rho2 > 1e-80 && th > 1e-18f) {
      const float ith = 1.f / th;
      float sh, ch;
      sincosf(0.5f * th, &sh, &ch);
      const float sn = 2.f * sh * ch, cc = 1.f - 2.f * sh * sh;
      const float q = sh * ith;                      /* w = (1 - cos th) / th^2 = 2 q^2 */
      const float w = 2.f * q * q, v = sn * ith;
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) R3[i][k] = (i == k ? cc : 0.f) + w * t[i] * t[k] + v * X[i][k];
    } else {
      /* the reference's small-rho form (below 1e-18 rad the rotation is this
       * to FP32's precision) */
      for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) R3[i][k] = (i == k ? 1.f : 0.f) + X[i][k];
    }
  }
  /* emission j x, scaled */
  const double jx[4] = {c.jI * x, c.jQ * x, c.jU * x, c.jV * x};
  double jmax = fmax(fmax(fabs(jx[0]), fabs(jx[1])), fmax(fabs(jx[2]), fabs(jx[3])));
  int e = 0;
  if (jmax > 0. && jmax < 1e300) (void)frexp(jmax, &e);
  const double down = ldexp(1., -e), up = ldexp(1., e);
  const float jn[4] = {(float)(jx[0] * down), (float)(jx[1] * down), (float)(jx[2] * down), (float)(jx[3] * down)};
  float A[4][4], b[4];
  for (int p = 0; p < 4; p++)
    for (int q = 0; q < 4; q++) A[p][q] = 0.f;
  {
    const double aP2 = c.aQ * c.aQ + c.aU * c.aU + c.aV * c.aV;
    if (aP2 > 1e-80) {
      const double aP = sqrt(aP2);
      const float tI = (float)(c.aI * x);
      const float dm = (float)((c.aI - aP) * x);      /* tau_I - tau_P */
      const float n[3] = {(float)(c.aQ / aP), (float)(c.aU / aP), (float)(c.aV / aP)};
      /* e^(-tau_I -+ tau_P) - 1 and e^(-tau_I) - 1 (expm1, so that thin
       * steps keep their relative accuracy) */
      const float tS = (float)((c.aI + aP) * x);      /* tau_I + tau_P */
      const float emm = expm1f(-dm), emp = expm1f(-tS), em = expm1f(-tI);
      const float efacm = 1.f + emm, efacp = 1.f + emp, efac = 1.f + em;
      const float ch = 0.5f * (efacm + efacp), sh = 0.5f * (efacp - efacm);
      A[0][0] = ch;
      for (int k = 0; k < 3; k++) {
        A[0][k + 1] = sh * n[k];
        A[k + 1][0] = sh * n[k];
        for (int l = 0; l < 3; l++) A[k + 1][l + 1] = (k == l ? efac : 0.f) + n[k] * n[l] * (ch - efac);
      }
      /* the source terms as (1 - e^-tau) / tau times scaled emission, so
       * that no product of two small optical depths is formed */
      const float nj = n[0] * jn[1] + n[1] * jn[2] + n[2] * jn[3];
      const float m = 0.5f * one_minus_exp_over(dm, emm) * (jn[0] - nj);
      const float pp = 0.5f * one_minus_exp_over(tS, emp) * (jn[0] + nj);
      const float pI = one_minus_exp_over(tI, em);
      b[0] = m + pp;
      for (int k = 0; k < 3; k++) b[k + 1] = pI * (jn[k + 1] - n[k] * nj) + (pp - m) * n[k];
    } else {
      const double dtau = x * c.aI;
      const float em = expm1f((float)(-dtau));
      const float g = one_minus_exp_over((float)dtau, em);
      for (int k = 0; k < 4; k++) {
        A[k][k] = 1.f + em;
        b[k] = jn[k] * g;
      }
    }
  }
  float R[4][4];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 4; k++) R[i][k] = (i == 0 || k == 0) ? (i == k ? 1.f : 0.f) : R3[i - 1][k - 1];
  float RA[4][4];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 4; k++) {
      float s = 0.f;
      for (int l = 0; l < 4; l++) s += R[i][l] * A[l][k];
      RA[i][k] = s;
    }
  for (int i = 0; i < 4; i++) {
    for (int k = 0; k < 4; k++) {
      float s = 0.f;
      for (int l = 0; l < 4; l++) s += RA[i][l] * R[l][k];
      T[i][k] = s;
    }
    float s = 0.f;
    for (int l = 0; l < 4; l++) s += R[i][l] * b[l];
    cv[i] = (double)s * up;
  }
}

/* Hermitian coherency tensor: n[0..3] real diagonal, then (re, im) of
 * N01 N02 N03 N12 N13 N23. */
struct Herm {
  double d[4];
  double re[6], im[6];
};

IPG_HD constexpr int off(int i, int j)  /* i < j */
{
  return (i == 0) ? (j - 1) : (i == 1) ? (j + 1) : 5;
}

IPG_HD void herm_get(const Herm &N, int i, int j, double &re, double &im)
{
  if (i == j) {
    re = N.d[i];
    im = 0.;
  } else if (i < j) {
    re = N.re[off(i, j)];
    im = N.im[off(i, j)];
  } else {
    re = N.re[off(j, i)];
    im = -N.im[off(j, i)];
  }
}

IPG_HD bool herm_nonzero(const Herm &N)
{
  bool z = true;
  for (int i = 0; i < 4; i++) z = z && (N.d[i] == 0.);
  for (int i = 0; i < 6; i++) z = z && (N.re[i] == 0.) && (N.im[i] == 0.);
  return !z;
}

/* Nout = Nin - dl (M Nm + Nm M^T): push_polar, src/ipolarray.c:253-289.
 * Each component accumulates as a chain of fused multiply-adds, and the
 * terms with the diagonal's zero imaginary part are left out (a product
 * with a zero operand cannot be dropped by the compiler under IEEE rules). */
IPG_HD void transport(const Herm &Nin, const Herm &Nm, const double M[4][4], double dl, Herm &Nout)
{
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = i; j < 4; j++) {
      double sr = 0., si = 0.;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        double ar, ai, br, bi;
        herm_get(Nm, k, j, ar, ai);
        herm_get(Nm, i, k, br, bi);
        sr = fma(M[i][k], ar, sr);
        sr = fma(M[j][k], br, sr);
        if (i != j) {
          if (k != j) si = fma(M[i][k], ai, si);
          if (k != i) si = fma(M[j][k], bi, si);
        }
      }
      if (i == j) {
        Nout.d[i] = fma(-dl, sr, Nin.d[i]);
      } else {
        Nout.re[off(i, j)] = fma(-dl, sr, Nin.re[off(i, j)]);
        Nout.im[off(i, j)] = fma(-dl, si, Nin.im[off(i, j)]);
      }
    }
}

/* Stokes parameters of N in a tetrad with covariant rows e1 = Ecov[1],
 * e2 = Ecov[2] (complex_coord_to_tetrad_rank2 + tensor_to_stokes,
 * src/ipolarray.c:684-694, :710-728, restricted to the [1..2] block):
 * T11 = e1.Re(N).e1, T22 = e2.Re(N).e2, Re T12 = e2.Re(N).e1 and
 * Im T12 = e1.Im(N).e2, as matrix-vector products in fused multiply-adds. */
IPG_HD void herm_to_stokes(const Herm &N, const double e1[4], const double e2[4], double S[4])
{
  double w1[4], w2[4], u2[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double a1 = N.d[k] * e1[k], a2 = N.d[k] * e2[k], b2 = 0.;
#pragma unroll
    for (int l = 0; l < 4; l++) {
      if (l == k) continue;
      double re, im;
      herm_get(N, k, l, re, im);
      a1 = fma(re, e1[l], a1);
      a2 = fma(re, e2[l], a2);
      b2 = fma(im, e2[l], b2);
    }
    w1[k] = a1;
    w2[k] = a2;
    u2[k] = b2;
  }
  double t11 = e1[0] * w1[0], t22 = e2[0] * w2[0], r12 = e2[0] * w1[0], i12 = e1[0] * u2[0];
#pragma unroll
  for (int k = 1; k < 4; k++) {
    t11 = fma(e1[k], w1[k], t11);
    t22 = fma(e2[k], w2[k], t22);
    r12 = fma(e2[k], w1[k], r12);
    i12 = fma(e1[k], u2[k], i12);
  }
  S[0] = (t11 + t22) / 2;
  S[1] = (t11 - t22) / 2;
  S[2] = r12;
  S[3] = -i12;
}

/* N += Econ^T dT Econ for a Stokes change dS in the [1..2] tetrad block
 * (stokes_to_tensor + complex_tetrad_to_coord_rank2, src/ipolarray.c:673-682,
 * :730-748, applied as a difference): with a = dI + dQ, b = dI - dQ, u = dU,
 * v = dV, Re dN_kl = p_k E1_l + q_k E2_l (p = a E1 + u E2, q = b E2 + u E1)
 * and Im dN_kl = v (E2_k E1_l - E1_k E2_l). */
IPG_HD void herm_add_stokes(Herm &N, const double E1[4], const double E2[4], const double dS[4])
{
  const double a = dS[0] + dS[1], b = dS[0] - dS[1], u = dS[2], v = dS[3];
  double pk[4], qk[4], vE1[4], vE2[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    pk[k] = fma(u, E2[k], a * E1[k]);
    qk[k] = fma(u, E1[k], b * E2[k]);
    vE1[k] = v * E1[k];
    vE2[k] = v * E2[k];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) N.d[k] = fma(pk[k], E1[k], fma(qk[k], E2[k], N.d[k]));
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int l = k + 1; l < 4; l++) {
      N.re[off(k, l)] = fma(pk[k], E1[l], fma(qk[k], E2[l], N.re[off(k, l)]));
      N.im[off(k, l)] = fma(vE2[k], E1[l], fma(-vE1[k], E2[l], N.im[off(k, l)]));
    }
}

/* Camera-frame Stokes of N, project_N (src/ipolarray.c:578-598). */
IPG_HD void project(const Herm &N, const double Ecam[4][4], const Metric &mcam, double rotcam,
                    double S[4])
{
  double e1[4], e2[4];
  lower(mcam.g, Ecam[1], e1);
  lower(mcam.g, Ecam[2], e2);
  double T[4];
  herm_to_stokes(N, e1, e2, T);
  double rc = -2. * rotcam;
  S[0] = T[0];
  S[1] = T[1] * cos(rc) - T[2] * sin(rc);
  S[2] = T[1] * sin(rc) + T[2] * cos(rc);
  S[3] = T[3];
}

}  // namespace ipg
