/*
 * Per-ray and per-step bodies of the pipeline, shared by the gfx950 kernels
 * in ipgpu.hip.
 *
 * trace_ray:  the backward geodesic of one pixel, src/model_geodesics.c:38-212
 *             with push_photon (src/geodesics.c:9-64).  Writes one record of
 *             REC doubles per step: X1 X2 X3 K0 K1 K2 K3 dl (code units).
 * prep_step:  the state-independent part of one transfer step of
 *             integrate_emission (src/ipolarray.c:59-247): connection
 *             contractions for parallel transport, fluid state, transfer
 *             coefficients and plasma tetrad.
 * chain_ray:  the state recurrence of integrate_emission over the packets
 *             of one ray, then project_N (:578-598).
 */
#pragma once

#include "ipg_physics.h"

namespace ipg {

constexpr int REC = 8;

/* keeps the scheduler from moving instructions across this point (bounds
 * the registers a long unrolled sequence holds) */
#if defined(__HIP_DEVICE_COMPILE__)
#define IPG_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)
#else
#define IPG_SCHED_BARRIER() ((void)0)
#endif

struct Rec {
  double X[4], K[4], dl;
};

IPG_HD void store_rec(double *o, const double X[4], const double K[4], double dl)
{
  o[0] = X[1];
  o[1] = X[2];
  o[2] = X[3];
  o[3] = K[0];
  o[4] = K[1];
  o[5] = K[2];
  o[6] = K[3];
  o[7] = dl;
}

IPG_HD void load_rec(const double *rec, int n, Rec &r)
{
  const double *p = rec + (size_t)n * REC;
#if defined(__HIP_DEVICE_COMPILE__)
  const double2 *q = reinterpret_cast<const double2 *>(p);
  double2 a = q[0], b = q[1], c = q[2], d = q[3];
  r.X[1] = a.x;
  r.X[2] = a.y;
  r.X[3] = b.x;
  r.K[0] = b.y;
  r.K[1] = c.x;
  r.K[2] = c.y;
  r.K[3] = d.x;
  r.dl = d.y;
#else
  r.X[1] = p[0];
  r.X[2] = p[1];
  r.X[3] = p[2];
  r.K[0] = p[3];
  r.K[1] = p[4];
  r.K[2] = p[5];
  r.K[3] = p[6];
  r.dl = p[7];
#endif
  r.X[0] = 0.;
}

/* Returns nstep.  Records beyond cap are not written (the caller retraces
 * with a larger cap; cap = 0 only counts the steps).  nt (may be null)
 * receives the subring count per step when P.target_nturns >= 0.
 * With nend, the integration pauses after nend steps (*finished = 0); with
 * n0 > 0 it resumes from record n0 (its X1..3 and K are the registers'
 * values after that step bit for bit; X0 is never read), without nt.
 *
 * push_photon (src/geodesics.c:15-49) in the order that shortens the
 * dependent chain of a step: the end point X' = X + dl Kh is known once
 * the half step is, so the metric at X' (the next step's start) is
 * evaluated together with the metric at the half-step point, two
 * independent straight-line evaluations the compiler interleaves.  The
 * values are those of the plain order. */
template <int MK, bool NT>
IPG_HD_INLINE int trace_ray_t(const ipgpu_geom &G, const ipgpu_image &P, double X[4], double K[4], int nstep,
                              int cap, double *rec, int *nt, int nend, int *finished)
{
  Metric m;
  int passed = (P.Xcam[2] < 0.5) ? 1 : ((P.Xcam[2] > 0.5) ? -1 : 0);
  int nturns = 0;
  double za = 0., zb = 0., x2a = 0., x2b = 0.;
  int done = 1;
  metric_k<MK>(G, X[1], X[2], m);
  while (true) {
    if (stop_backward(G, m.r, K)) break;
    if (nstep >= P.maxnstep - 1) break;
    if (nstep >= nend) {
      done = 0;
      break;
    }
    nstep++;
    double dl = stepsize(G, X, K, P.eps);
    double dlp = -dl;
    double Xh[4], Kh[4], Xn[4];
    half_step(m, X, K, dlp, Xh, Kh);
    for (int k = 0; k < 4; k++) Xn[k] = X[k] + dlp * Kh[k];
    Metric mh;
    metric_k<MK>(G, Xh[1], Xh[2], mh);
    metric_k<MK>(G, Xn[1], Xn[2], m);
    double acc[4];
    geodesic_accel(mh, Kh, acc);
    for (int k = 0; k < 4; k++) K[k] = K[k] - dlp * acc[k];
    for (int k = 0; k < 4; k++) X[k] = Xn[k];
    if (nstep < cap) store_rec(rec + (size_t)nstep * REC, X, K, dl);
    if constexpr (NT) {
      /* subring turning points, src/model_geodesics.c:129-165 */
      double rr, th;
      bl_coord(G, X[1], X[2], rr, th);
      double zc = rr * cos(th), x2c = X[2];
      if (passed != 0) {
        if (passed > 0) {
          if (X[2] > 0.5) passed = 0;
        } else {
          if (X[2] < 0.5) passed = 0;
        }
      } else if (nstep > 3) {
        if (P.subring_dtheta == 1) {
          if ((x2b - x2a) * (x2c - x2b) < 0) nturns++;
        } else {
          if ((zb - za) * (zc - zb) < 0) nturns++;
        }
      }
      za = zb;
      zb = zc;
      x2a = x2b;
      x2b = x2c;
      if (nstep < cap) nt[nstep] = nturns;
    }
  }
  if (finished) *finished = done;
  return nstep;
}

IPG_HD int trace_ray(const ipgpu_geom &G, const ipgpu_image &P, int pi, int pj, int cap, double *rec,
                     int *nt, int n0 = 0, int nend = 0x7fffffff, int *finished = nullptr)
{
  double X[4], K[4];
  if (n0 > 0) {
    Rec r;
    load_rec(rec, n0, r);
    X[0] = 0.;
    for (int k = 1; k < 4; k++) X[k] = r.X[k];
    for (int k = 0; k < 4; k++) K[k] = r.K[k];
  } else {
    double E[4][4];
    Metric m;
    for (int k = 0; k < 4; k++) X[k] = P.Xcam[k];
    camera_tetrad(G, X, P.old_centering, E, m);
    init_K(P, E, pi, pj, K);
    if (cap > 0) store_rec(rec, X, K, 0.);
    if (nt && cap > 0) nt[0] = 0;
    n0 = 0;
  }
#define IPG_TRACE_CASE(MKV)                                                          \
  case MKV:                                                                           \
    return nt ? trace_ray_t<MKV, true>(G, P, X, K, n0, cap, rec, nt, nend, finished)   \
              : trace_ray_t<MKV, false>(G, P, X, K, n0, cap, rec, nt, nend, finished);
  switch (metric_kind(G)) {
    IPG_TRACE_CASE(MK_MINK)
    IPG_TRACE_CASE(MK_MKS)
    IPG_TRACE_CASE(MK_FMKS_INT)
    default: IPG_TRACE_CASE(MK_FMKS)
  }
#undef IPG_TRACE_CASE
}

/* Step packet: everything one transfer step needs that does not depend on
 * the state N, I of the ray, so that it can be formed for all steps of all
 * rays in parallel (prep_step) and consumed by a short serial recurrence
 * (chain_ray).  Layout (doubles):
 *   0 meta      bits: 1 radiating, 2 zero_emission, 4..: oddflag << 2
 *   1 dl_cgs    2 ju, 3 ku (unpolarized j/nu^2, a nu at the step end point)
 *   4 dl/2, 5 dl (code units, for push_polar)   6 dl_cgs |rho_V|   7 unused
 *   8..23 T, 24..27 cv   the plasma-frame Stokes step S -> T S + cv
 *   28..31 E1, 32..35 E2   plasma tetrad e^(1), e^(2) (contravariant)
 *   36..39 e1, 40..43 e2   the same lowered
 *   44..59 Mh   Gamma^i_kl K^l at the step midpoint
 *   60..75 Mf   Gamma^i_kl K^l at the step end point                    */
constexpr int PK = 76;      /* polarized packet */
constexpr int PK_UNPOL = 4; /* -unpol: meta, dl, ju, ku only */
constexpr int PK_META = 0, PK_DL = 1, PK_JU = 2, PK_KU = 3, PK_HDL = 4, PK_FDL = 5, PK_DTF = 6,
              PK_T = 8, PK_CV = 24, PK_E1 = 28, PK_E2 = 32, PK_e1 = 36, PK_e2 = 40, PK_MH = 44,
              PK_MF = 60;

/* Rescaling of a stored trajectory and of the dump for another job of a
 * batch.  The geodesic in X does not depend on the frequency: K scales with
 * it and the affine step dl inversely (stepsize, src/model_geodesics.c:
 * 267-293, is homogeneous of degree -1 in K), so one trace serves every
 * frequency of a camera.  n_e scales with M_unit and B with its square
 * root (set_units, model/iharm/model.c:595-608); Theta_e, sigma and beta do
 * not depend on it.  {1, 1, 1, 1} leaves everything bit for bit unchanged. */
struct Scale {
  double kf, ikf, ne, b;
};
IPG_HD Scale unit_scale()
{
  Scale s;
  s.kf = s.ikf = s.ne = s.b = 1.;
  return s;
}

IPG_HD void scale_rec(Rec &r, const Scale &sc)
{
  for (int k = 0; k < 4; k++) r.K[k] *= sc.kf;
  r.dl *= sc.ikf;
}

/* Packet of transfer step n (from trajectory point n to n-1) of
 * integrate_emission, src/ipolarray.c:78-180.  nt may be null. */
/* Adjoint record of one step (AX doubles), written by prep_step when asked:
 * the inputs of the coefficients (fluid-frame frequency and angle, n_e,
 * Theta_e, B) and of their dependence on the dump (interpolation cell and
 * weights, sigma smoothing, dTheta_e/dR_high, dR_low).  ax[4] = 0 where
 * the step does not radiate or n_e = 0. */
/* adjoint record of a step: the coefficient map's inputs, the interpolation
 * cell and weights, dTheta_e/dR_high, dTheta_e/dR_low, and the derivatives
 * of the step's affine Stokes map S -> T S + cv with respect to p = ln n_e,
 * Theta_e, ln B: dT/dp at AX_DT + 16 p (row-major), dcv/dp at AX_DCV + 4 p */
constexpr int AX = 76;
constexpr int AX_NU = 0, AX_STH = 1, AX_CTH = 2, AX_EDGE = 3, AX_NE = 4, AX_TH = 5, AX_B = 6, AX_CELL = 7,
              AX_D1 = 8, AX_D2 = 9, AX_D3 = 10, AX_SMOOTH = 11, AX_DRH = 12, AX_DRL = 13, AX_ZERO = 14,
              AX_DT = 16, AX_DCV = 64;

/* Position of field f in a packet stored with field stride PS.  PS = 1:
 * one packet contiguous.  PS = 64: the packets of the 64 rays of a wave
 * interleaved by field pairs, fields 2m and 2m + 1 of lane l's packet at
 * 128 m + 2 l and + 1 (the caller offsets lane l's packet by 2 l), so that
 * a wave's load of a pair is one 16-byte access per lane, 1 KiB contiguous,
 * and a packet 38 loads: few enough that the next step's packet stays in
 * flight under the hardware's count of outstanding loads. */
template <int PS>
IPG_HD constexpr size_t pk_at(int f)
{
  return PS == 1 ? (size_t)f : (size_t)(f >> 1) * (2 * PS) + (f & 1);
}

/* Lane l's offset within a PS-interleaved packet row. */
template <int PS>
IPG_HD constexpr size_t pk_lane(int l)
{
  return PS == 1 ? 0 : (size_t)2 * l;
}

/* Packet fields are written through PkW: field f of the packet at p is
 * p[pk_at<PS>(f)]. */
template <int PS>
struct PkW {
  double *p;
  IPG_HD double &operator[](int f) const { return p[pk_at<PS>(f)]; }
};

#if !defined(IPG_OPCOUNT)
/* Derivatives of the step's affine Stokes map S -> T S + cv (stokes_affine
 * of the coefficients at the step) with respect to p = ln n_e, Theta_e and
 * ln B, by central differences of the coefficient map (relative step
 * 1e-5), from the inputs the record ax holds; written to ax[AX_DT + 16 p +
 * 4 i + j] and ax[AX_DCV + 4 p + i]. */
IPG_HD void step_param_maps(double *ax, int etype, double dl)
{
  Geo q;
  q.nu = ax[AX_NU];
  q.sth = ax[AX_STH];
  q.cth = ax[AX_CTH];
  q.edge = ax[AX_EDGE] != 0.;
  const double Ne = ax[AX_NE], Th = ax[AX_TH], B = ax[AX_B];
  const bool zero = ax[AX_ZERO] != 0.;
  const double eps = 1.e-5;
  for (int p = 0; p < 3; p++) {
    /* ln(1+eps) - ln(1-eps) for the logarithmic directions, 2 eps Theta_e for Theta_e */
    const double inv = 1. / ((p == 1) ? 2. * eps * Th : (log(1. + eps) - log(1. - eps)));
    double *dT = ax + AX_DT + 16 * p, *dc = ax + AX_DCV + 4 * p;
    for (int sgn = 0; sgn < 2; sgn++) {
      const double f = sgn ? (1. - eps) : (1. + eps);
      double ne = Ne, th = Th, b = B;
      if (p == 0) ne *= f;
      else if (p == 1) th *= f;
      else b *= f;
      double ju, ku;
      Coef c;
      coefficients_core(q, ne, th, b, etype, true, ju, ku, c);
      if (zero) c.jI = c.jQ = c.jU = c.jV = 0.;
      double T[4][4], cv[4];
      stokes_affine(c, dl, T, cv);
      if (sgn == 0) {
        for (int i = 0; i < 4; i++) {
          for (int j = 0; j < 4; j++) dT[4 * i + j] = T[i][j];
          dc[i] = cv[i];
        }
      } else {
        for (int i = 0; i < 4; i++) {
          for (int j = 0; j < 4; j++) dT[4 * i + j] = (dT[4 * i + j] - T[i][j]) * inv;
          dc[i] = (dc[i] - cv[i]) * inv;
        }
      }
    }
  }
}
#endif

/* The source of the step's records and interpolated dump fields when the
 * caller has loaded or staged them (records / fetch fill their arguments
 * and return true); NoVG: prep_step loads the records and gathers the
 * fields from the dump itself. */
struct NoVG {
  IPG_HD bool records(Rec &, Rec &) const { return false; }
  IPG_HD bool fetch(double *) const { return false; }
};

template <bool F32 = false, int PS = 1, class VG = NoVG>
IPG_HD_INLINE void prep_step(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P, int n,
                      const double *rec, const int *nt, const double *dump, double *pk_out, bool pol,
                      const Scale &sc, double *ax = nullptr, const VG &vg = VG(), const float *d32 = nullptr)
{
  PkW<PS> pk{pk_out};
  if (ax)
    for (int q = 0; q < AX; q++) ax[q] = 0.;
  Rec ti, tf;
  if (!vg.records(ti, tf)) {
    load_rec(rec, n, ti);
    load_rec(rec, n - 1, tf);
  }
  scale_rec(ti, sc);
  scale_rec(tf, sc);
  const double conv = Mo.L_unit * HPL / (ME * CL * CL);
  const double dl_cgs = ti.dl * conv;
  Metric m;
  metric(G, tf.X[1], tf.X[2], m);
  /* staged fields are read here, before the packet stores begin */
  double vst[10];
  const bool staged = vg.fetch(vst);
  int meta = 0;
  pk[PK_DL] = dl_cgs;
  if (pol) {
    /* each transport matrix stored as soon as it is formed: the step-end
     * matrix is not held while the half-step metric is evaluated */
    double Xh[4], Kh[4];
    {
      double Mf[4][4];
      transport_matrix(m, tf.K, Mf);
      for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) pk[PK_MF + 4 * i + k] = Mf[i][k];
    }
    half_step(m, tf.X, tf.K, -ti.dl, Xh, Kh);
    {
      Metric mh;
      metric(G, Xh[1], Xh[2], mh);
      double Mh[4][4];
      transport_matrix(mh, Kh, Mh);
      for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) pk[PK_MH + 4 * i + k] = Mh[i][k];
    }
    pk[PK_HDL] = (0.5 * dl_cgs) / conv;
    pk[PK_FDL] = dl_cgs / conv;
    pk[PK_DTF] = 0.;
    pk[7] = 0.;
  }
  if (radiating_region(G, Mo, m.r, m.th)) {
    meta |= 1;
    int zero_emission = 0;
    if (P.target_nturns >= 0 && nt && nt[n] != P.target_nturns) zero_emission = 1;
    const double half_x2 = (G.cstopx[2] - G.cstartx[2]) / 2;
    if (P.counterjet == 1) {
      if (tf.X[2] < half_x2) zero_emission = 1;
    } else if (P.counterjet == 2) {
      if (tf.X[2] > half_x2) zero_emission = 1;
    }
    if (zero_emission) meta |= 2;
    Fluid F;
    FluidExt ext;
    /* mixed precision with d32: the corners from the FP32 copy of the records */
    bool pre = staged;
    if (F32 && !staged && d32 && Mo.kind != IPGPU_MODEL_SPHERE && in_domain(G, tf.X[1], tf.X[2])) {
      Cell c32;
      locate(G, tf.X[1], tf.X[2], tf.X[3], c32);
      gather10_f32(G, d32, c32, vst);
      pre = true;
    }
    fluid(G, Mo, dump, tf.X, m, F, ax ? &ext : nullptr, vst, pre);
    F.ne *= sc.ne;
    F.B *= sc.b;
    if (ax && F.ne > 0.) {
      Geo q = coef_geometry(F, tf.K);
      ax[AX_NU] = q.nu;
      ax[AX_STH] = q.sth;
      ax[AX_CTH] = q.cth;
      ax[AX_EDGE] = q.edge ? 1. : 0.;
      ax[AX_NE] = F.ne;
      ax[AX_TH] = F.thetae;
      ax[AX_B] = F.B;
      ax[AX_CELL] = ext.cell;
      ax[AX_D1] = ext.d1;
      ax[AX_D2] = ext.d2;
      ax[AX_D3] = ext.d3;
      ax[AX_SMOOTH] = ext.smooth;
      ax[AX_DRH] = ext.dth_rh;
      ax[AX_DRL] = ext.dth_rl;
      ax[AX_ZERO] = zero_emission ? 1. : 0.;
#if !defined(IPG_OPCOUNT)
      if (pol) step_param_maps(ax, P.emission_type, dl_cgs);
#endif
    }
    Coef c;
    double ju, ku;
    if (F32) coefficients_f32(F, tf.K, P.emission_type, pol, ju, ku, c);
    else coefficients(F, tf.K, P.emission_type, pol, ju, ku, c);
    pk[PK_JU] = ju;
    pk[PK_KU] = ku;
    if (pol) {
      if (zero_emission) c.jI = c.jQ = c.jU = c.jV = 0.;
      double T[4][4], cv[4];
      if (F32) stokes_affine_f32(c, dl_cgs, T, cv);
      else stokes_affine(c, dl_cgs, T, cv);
      for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 4; k++) pk[PK_T + 4 * i + k] = T[i][k];
        pk[PK_CV + i] = cv[i];
      }
      pk[PK_DTF] = dl_cgs * fabs(c.rV);
      /* plasma tetrad of evolve_N, src/ipolarray.c:370-384 */
      double Bcon[4] = {F.Bcon[0], F.Bcon[1], F.Bcon[2], F.Bcon[3]};
      double bsq = F.Bcon[0] * F.Bcov[0] + F.Bcon[1] * F.Bcov[1] + F.Bcon[2] * F.Bcov[2]
          + F.Bcon[3] * F.Bcov[3];
      if (bsq <= 0.) {
        Bcon[0] = 0.;
        Bcon[1] = 1.;
        Bcon[2] = 1.;
        Bcon[3] = 1.;
      }
      double E[4][4];
      int flag = plasma_tetrad(F.Ucon, tf.K, Bcon, m.g, m.sqrtmg, E);
      if (isnan(F.Ucov[0])) flag |= 8;
      double e1[4], e2[4];
      lower(m.g, E[1], e1);
      lower(m.g, E[2], e2);
      for (int q = 0; q < 4; q++) {
        pk[PK_E1 + q] = E[1][q];
        pk[PK_E2 + q] = E[2][q];
        pk[PK_e1 + q] = e1[q];
        pk[PK_e2 + q] = e2[q];
      }
      meta |= flag << 2;
    }
  }
  pk[PK_META] = (double)meta;
}

/* Unpolarized coefficients at the far end of the trajectory (the cached
 * js, ks of integrate_emission, src/ipolarray.c:74). */
template <bool F32 = false>
IPG_HD void far_end_coefficients(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P,
                                 int ns, const double *rec, const double *dump, double &js, double &ks,
                                 const Scale &sc)
{
  Rec t;
  load_rec(rec, ns, t);
  scale_rec(t, sc);
  Metric m;
  metric(G, t.X[1], t.X[2], m);
  Fluid F;
  fluid(G, Mo, dump, t.X, m, F);
  F.ne *= sc.ne;
  F.B *= sc.b;
  Coef c;
  if (F32) coefficients_f32(F, t.K, P.emission_type, false, js, ks, c);
  else coefficients(F, t.K, P.emission_type, false, js, ks, c);
}

/* One packet held in registers; on the device it is read with 16-byte
 * loads so that the next step's packet is in flight while the current step
 * computes. */
struct Pkt {
  double v[PK];
};

/* A packet stored with field stride PS (see pk_at), one field pair per
 * load. */
template <int PS>
IPG_HD_INLINE void load_pkt_s(const double *p, Pkt &k)
{
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
  for (int i = 0; i < PK / 2; i++) {
    const double2 t = *reinterpret_cast<const double2 *>(p + pk_at<PS>(2 * i));
    k.v[2 * i] = t.x;
    k.v[2 * i + 1] = t.y;
  }
#else
  for (int i = 0; i < PK; i++) k.v[i] = p[pk_at<PS>(i)];
#endif
}

IPG_HD void load_pkt(const double *p, int n, Pkt &k)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
  for (int i = 0; i < PK / 2; i++) {
    double2 t = q[i];
    k.v[2 * i] = t.x;
    k.v[2 * i + 1] = t.y;
  }
#else
  for (int i = 0; i < n; i++) k.v[i] = p[i];
#endif
}

/* S0 = Stokes(N) in the plasma frame, S = T S0 + cv, optional floors, and
 * N += Econ^T (S - S0) Econ: the state part of evolve_N
 * (src/ipolarray.c:386-575).  Returns oddflag bits 2, 4. */
IPG_HD int stokes_apply(Herm &N, const double *pk, int stokes_floors, double &tauF)
{
  double S0[4], S[4];
  herm_to_stokes(N, pk + PK_e1, pk + PK_e2, S0);
  for (int i = 0; i < 4; i++)
    S[i] = pk[PK_CV + i] + pk[PK_T + 4 * i] * S0[0] + pk[PK_T + 4 * i + 1] * S0[1]
        + pk[PK_T + 4 * i + 2] * S0[2] + pk[PK_T + 4 * i + 3] * S0[3];
  tauF += pk[PK_DTF];
  if (stokes_floors) {
    if (S[0] < 0) {
      S[0] = S[1] = S[2] = S[3] = 0.;
    } else {
      double pf = sqrt(S[1] * S[1] + S[2] * S[2] + S[3] * S[3]) / S[0];
      if (pf > 1.) {
        S[1] /= pf;
        S[2] /= pf;
        S[3] /= pf;
      }
    }
  }
  double dS[4] = {S[0] - S0[0], S[1] - S0[1], S[2] - S0[2], S[3] - S0[3]};
  herm_add_stokes(N, pk + PK_E1, pk + PK_E2, dS);
  int flag = 0;
  if (tauF > 1.e100 || tauF < -1.e100 || isnan(tauF)) flag |= 2;
  if (isnan(N.d[0]) || isnan(S[0])) flag |= 4;
  return flag;
}

/* State of the transfer recurrence of one ray. */
struct Chain {
  Herm N;
  double Mi[4][4];
  double I, tau, tauF, js, ks;
  int flag, nrad;
  bool Nnz;
};

IPG_HD void chain_init(Chain &st, double js, double ks)
{
  for (int q = 0; q < 4; q++) st.N.d[q] = 0.;
  for (int q = 0; q < 6; q++) st.N.re[q] = st.N.im[q] = 0.;
  st.I = st.tau = st.tauF = 0.;
  st.js = js;
  st.ks = ks;
  st.flag = st.nrad = 0;
  st.Nnz = false;
}

/* One step of integrate_emission's loop body (src/ipolarray.c:78-180) on
 * the state, from the step's packet. */
IPG_HD_INLINE void chain_step(Chain &st, const double *pk, bool pol, int stokes_floors)
{
  const int meta = (int)pk[PK_META];
  if (pol && st.Nnz) {
    /* push_polar twice, src/ipolarray.c:85-89 */
    double Mh[4][4];
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 4; k++) Mh[i][k] = pk[PK_MH + 4 * i + k];
    Herm Nh;
    transport(st.N, st.N, st.Mi, pk[PK_HDL], Nh);
    transport(st.N, Nh, Mh, pk[PK_FDL], st.N);
  }
  if (meta & 1) {
    st.nrad++;
    double jf = pk[PK_JU], kf = pk[PK_KU];
    double ji = (meta & 2) ? 0. : st.js, ki = st.ks;
    st.js = jf;
    st.ks = kf;
    st.I = approximate_solve(st.I, ji, ki, (meta & 2) ? 0. : jf, kf, pk[PK_DL], st.tau);
    if (pol) {
      st.flag |= stokes_apply(st.N, pk, stokes_floors, st.tauF);
      st.flag |= meta >> 2;
      st.Nnz = herm_nonzero(st.N);
    }
  }
  if (pol && st.Nnz)
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 4; k++) st.Mi[i][k] = pk[PK_MF + 4 * i + k];
}

/* project_N (src/ipolarray.c:578-598) and the per-ray outputs:
 * out7 = Intensity, Is, Qs, Us, Vs, Tau, tauF (before the nu^3 scaling of
 * save_pixel); returns the oddflag bits. */
IPG_HD int chain_finish(const ipgpu_geom &G, const ipgpu_image &P, const Chain &st, double out7[7])
{
  double S[4] = {0., 0., 0., 0.};
  if (!P.only_unpolarized) {
    double E[4][4];
    Metric m;
    const double Xc[4] = {P.Xcam[0], P.Xcam[1], P.Xcam[2], P.Xcam[3]};
    camera_tetrad(G, Xc, 0, E, m);
    project(st.N, E, m, P.rotcam, S);
  }
  out7[0] = st.I;
  out7[1] = S[0];
  out7[2] = S[1];
  out7[3] = S[2];
  out7[4] = S[3];
  out7[5] = st.tau;
  out7[6] = st.tauF;
  return st.flag;
}

/* The state recurrence over the packets of one ray held in memory (packet
 * 0 is step ns, packet ns-1 is step 1), then project_N.  On the device the
 * next packet is loaded while the current step computes. */
template <int PS = 1>
IPG_HD_INLINE int chain_ray(const ipgpu_geom &G, const ipgpu_image &P, int ns, const double *pkt, int stride,
                     double js, double ks, double out7[7], int &nrad)
{
  const bool pol = stride == PK;
  Chain st;
  chain_init(st, js, ks);
  if (pol) {
    /* two packet buffers used in turn (no copy between them): the next
     * step's packet is in flight while the current one computes */
    Pkt a, b;
    auto load = [&](int q, Pkt &k) {
      if constexpr (PS == 1) load_pkt(pkt + (size_t)q * PK, PK, k);
      else load_pkt_s<PS>(pkt + (size_t)q * PK * PS, k);
    };
    /* the loads are unconditional (past the last step they re-read the
     * last packet), so that the compiler's count of loads in flight is the
     * same on every path and the current step waits for its own packet only */
    if (ns > 0) load(0, a);
    for (int q = 0; q < ns; q += 2) {
      load(q + 1 < ns ? q + 1 : ns - 1, b);
      chain_step(st, a.v, true, P.stokes_floors);
      if (q + 1 >= ns) break;
      load(q + 2 < ns ? q + 2 : ns - 1, a);
      chain_step(st, b.v, true, P.stokes_floors);
    }
  } else {
    for (int q = 0; q < ns; q++) {
      const double *p = pkt + (size_t)q * stride * PS;
      double u[PK_UNPOL];
      for (int f = 0; f < PK_UNPOL; f++) u[f] = p[pk_at<PS>(f)];
      chain_step(st, u, false, P.stokes_floors);
    }
  }
  nrad = st.nrad;
  return chain_finish(G, P, st, out7);
}

/* ------------------------------------------- long rays in segments
 *
 * The polarized recurrence is affine in the 16 real parameters of N
 * (d[4], re[6], im[6]): each step maps N -> A_q N + b_q (transport, then the
 * plasma-frame Stokes step; the transport of N = 0 is 0, so transporting
 * unconditionally is the same map).  The unpolarized recurrence is affine
 * in I (approximate_solve: I -> a I + b, with a, b functions of the
 * coefficients), and tau, tauF are sums.  A long ray's steps are cut into
 * segments of seg_len steps whose composed maps are formed in parallel
 * (16 columns of the N map from the basis states without sources, one from
 * N = 0 with sources, one column for I, tau, tauF), then applied in order
 * from the far end: the same result up to the roundoff of the regrouping. */
constexpr int SEG_LEN = 384;          /* default segment length */
constexpr int SEG_MIN_STEPS = 3584;   /* default: rays longer than this use segments */
constexpr int NPAR = 16;
constexpr int SEG_I = NPAR + 1;       /* column of the unpolarized map */
constexpr int SEGCOL = NPAR + 2;
/* slots of the unpolarized column: the map I -> a I + b of the segment's
 * radiating steps after its first one (which enters with the emissivity and
 * absorptivity carried from earlier segments and is applied when the
 * segments are composed), their tau, the segment's tauF and radiating step
 * count, the first radiating step's data and the last one's ju, ku (the
 * carry out) */
constexpr int SI_A = 0, SI_B = 1, SI_TAU = 2, SI_TAUF = 3, SI_NRAD = 4, SI_M2 = 5, SI_J1 = 6, SI_K1 = 7,
              SI_DL1 = 8, SI_JL = 9, SI_KL = 10;

/* Segment sg of a chain of ns steps: packets [q0, q1).  The boundaries sit
 * at trace steps that are multiples of seg_len counted from the camera (the
 * far-end segment is the short one), so that the segments of a ray's first
 * steps do not depend on its length. */
IPG_HD void segment_bounds(int ns, int seg_len, int sg, int &q0, int &q1)
{
  const int nseg = (ns + seg_len - 1) / seg_len;
  const int f = ns - (nseg - 1) * seg_len;
  q0 = sg == 0 ? 0 : f + (sg - 1) * seg_len;
  q1 = f + sg * seg_len;
}

IPG_HD double &herm_par(Herm &N, int b)
{
  return b < 4 ? N.d[b] : (b < 10 ? N.re[b - 4] : N.im[b - 10]);
}

IPG_HD double herm_par(const Herm &N, int b)
{
  return b < 4 ? N.d[b] : (b < 10 ? N.re[b - 4] : N.im[b - 10]);
}

/* One step of a segment column on N from the step's packet v (fields by
 * index): transport with the previous step's end-point matrix Mi (when
 * transport_on), then the plasma-frame Stokes step with its source term
 * when src. */
IPG_HD_INLINE void segment_step(Herm &N, const double *v, const double Mi[4][4], bool transport_on, bool src,
                                int &flag)
{
  if (transport_on) {
    double Mh[4][4];
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 4; k++) Mh[i][k] = v[PK_MH + 4 * i + k];
    Herm Nh;
    transport(N, N, Mi, v[PK_HDL], Nh);
    transport(N, Nh, Mh, v[PK_FDL], N);
  }
  const int meta = (int)v[PK_META];
  if (meta & 1) {
    double S0[4], dS[4], e1[4], e2[4], E1[4], E2[4];
    for (int i = 0; i < 4; i++) {
      e1[i] = v[PK_e1 + i];
      e2[i] = v[PK_e2 + i];
      E1[i] = v[PK_E1 + i];
      E2[i] = v[PK_E2 + i];
    }
    herm_to_stokes(N, e1, e2, S0);
    for (int i = 0; i < 4; i++)
      dS[i] = (src ? v[PK_CV + i] : 0.) + v[PK_T + 4 * i] * S0[0] + v[PK_T + 4 * i + 1] * S0[1]
          + v[PK_T + 4 * i + 2] * S0[2] + v[PK_T + 4 * i + 3] * S0[3] - S0[i];
    herm_add_stokes(N, E1, E2, dS);
    flag |= meta >> 2;
  }
}

/* Column b of the map of steps [q0, q1): applied to basis state b (b < 16,
 * no sources) or to N = 0 with sources (b = 16).  Returns the steps'
 * tetrad oddflag bits.  Two packet buffers used in turn, as in chain_ray:
 * the next step's packet is in flight while the current one computes (past
 * the segment's last step the loads re-read its packet). */
template <int PS = 1>
IPG_HD int segment_column(const double *pkt, int q0, int q1, int b, Herm &N)
{
  /* the basis state with constant indices only (a runtime index into N
   * would put it in scratch memory) */
#pragma unroll
  for (int q = 0; q < 4; q++) N.d[q] = (b == q) ? 1. : 0.;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    N.re[q] = (b == 4 + q) ? 1. : 0.;
    N.im[q] = (b == 10 + q) ? 1. : 0.;
  }
  const bool src = (b == NPAR);
  int flag = 0;
  /* end-point matrix of the step before the current one */
  double Mi[4][4];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 4; k++)
      Mi[i][k] = q0 > 0 ? pkt[(size_t)(q0 - 1) * PK * PS + pk_at<PS>(PK_MF + 4 * i + k)] : 0.;
  auto load = [&](int q, Pkt &k) {
    if constexpr (PS == 1) load_pkt(pkt + (size_t)q * PK, PK, k);
    else load_pkt_s<PS>(pkt + (size_t)q * PK * PS, k);
  };
  auto step = [&](int q, const Pkt &k) {
    segment_step(N, k.v, Mi, q > 0, src, flag);
    for (int i = 0; i < 4; i++)
      for (int c = 0; c < 4; c++) Mi[i][c] = k.v[PK_MF + 4 * i + c];
  };
  if (q1 > q0) {
    Pkt a, c;
    load(q0, a);
    for (int q = q0; q < q1; q += 2) {
      load(q + 1 < q1 ? q + 1 : q1 - 1, c);
      step(q, a);
      if (q + 1 >= q1) break;
      load(q + 2 < q1 ? q + 2 : q1 - 1, a);
      step(q + 1, c);
    }
  }
  return flag;
}

/* approximate_solve as the affine map I -> a I + b it is (the same branch
 * on dtau; the regrouping moves roundoff only). */
IPG_HD void approximate_solve_affine(double ji, double ki, double jf, double kf, double dl, double &tau,
                                     double &a, double &b)
{
  double javg = (ji + jf) / 2.;
  double kavg = (ki + kf) / 2.;
  double dtau = dl * kavg;
  tau += dtau;
  if (dtau < 1.e-3) {
    double f = dl * (1. - (dtau * 0.5) * (1. - dtau * (1. / 3.)));
    a = 1. - kavg * f;
    b = javg * f;
    return;
  }
  double efac = exp_neg(-dtau);
  a = efac;
  b = (javg * rcp(kavg)) * (1. - efac);
}

/* The unpolarized column of the map of steps [q0, q1) (slots SI_*).  The
 * recurrence caches the emissivity and absorptivity of the last radiating
 * step (src/ipolarray.c:140-146); within the segment that cache is local
 * except at its first radiating step. */
template <int PS = 1>
IPG_HD void segment_intensity(const double *pkt, int q0, int q1, double *o)
{
  double A = 1., B = 0., tau = 0., tauF = 0., js = 0., ks = 0.;
  int nrad = 0;
  for (int q = 0; q < NPAR; q++) o[q] = 0.;
  /* the five fields a step reads (meta, dl, ju, ku, dtauF), the next
   * step's in flight while the current one computes */
  auto load = [&](int q, double u[5]) {
    const double *pk = pkt + (size_t)q * PK * PS;
    u[0] = pk[pk_at<PS>(PK_META)];
    u[1] = pk[pk_at<PS>(PK_DL)];
    u[2] = pk[pk_at<PS>(PK_JU)];
    u[3] = pk[pk_at<PS>(PK_KU)];
    u[4] = pk[pk_at<PS>(PK_DTF)];
  };
  double cur[5], nxt[5];
  if (q1 > q0) load(q0, cur);
  for (int q = q0; q < q1; q++) {
    load(q + 1 < q1 ? q + 1 : q1 - 1, nxt);
    const int meta = (int)cur[0];
    if (meta & 1) {
      const double ju = cur[2], ku = cur[3];
      tauF += cur[4];
      if (nrad == 0) {
        o[SI_M2] = (meta & 2) ? 1. : 0.;
        o[SI_J1] = ju;
        o[SI_K1] = ku;
        o[SI_DL1] = cur[1];
      } else {
        double a, b;
        approximate_solve_affine((meta & 2) ? 0. : js, ks, (meta & 2) ? 0. : ju, ku, cur[1], tau, a, b);
        A = a * A;
        B = a * B + b;
      }
      js = ju;
      ks = ku;
      nrad++;
    }
    for (int f = 0; f < 5; f++) cur[f] = nxt[f];
  }
  o[SI_A] = A;
  o[SI_B] = B;
  o[SI_TAU] = tau;
  o[SI_TAUF] = tauF;
  o[SI_NRAD] = (double)nrad;
  o[SI_JL] = js;
  o[SI_KL] = ks;
}

/* State of the composition of a long ray's segment maps. */
struct SegChain {
  Herm N;
  double I, tau, tauF, js, ks;
  int flag, nrad;
};

/* Applies one segment's map A (SEGCOL x NPAR doubles, column-major, the
 * last column the unpolarized map) with oddflag bits fl.  The unpolarized
 * part comes first: after the N map, its branch kept the map's loads and
 * the products live across it (a 450-byte spill per lane). */
IPG_HD_INLINE void compose_segment(SegChain &c, const double *A, int fl)
{
  const double *u = A + SEG_I * NPAR;
  const int nr = (int)u[SI_NRAD];
  if (nr > 0) {
    /* the segment's first radiating step with the carried js, ks, as the
     * serial recurrence does it, then the rest of the segment */
    const bool m2 = u[SI_M2] != 0.;
    c.I = approximate_solve(c.I, m2 ? 0. : c.js, c.ks, m2 ? 0. : u[SI_J1], u[SI_K1], u[SI_DL1], c.tau);
    c.I = u[SI_A] * c.I + u[SI_B];
    c.tau += u[SI_TAU];
    c.js = u[SI_JL];
    c.ks = u[SI_KL];
  }
  c.tauF += u[SI_TAUF];
  c.nrad += nr;
  c.flag |= fl;
  double x[NPAR], acc[NPAR];
#pragma unroll
  for (int b = 0; b < NPAR; b++) x[b] = herm_par(c.N, b);
#pragma unroll
  for (int p = 0; p < NPAR; p++) acc[p] = A[NPAR * NPAR + p];
  /* column by column (acc[p] sums in the order b = 0..15), one column's
   * loads at a time so that the map is never held in registers whole */
#pragma unroll
  for (int b = 0; b < NPAR; b++) {
#pragma unroll
    for (int p = 0; p < NPAR; p++) acc[p] += A[b * NPAR + p] * x[b];
    IPG_SCHED_BARRIER();
  }
#pragma unroll
  for (int p = 0; p < NPAR; p++) herm_par(c.N, p) = acc[p];
}

/* A long ray from its segment maps (segmap: SEGCOL x NPAR doubles per
 * segment; segflag: the segments' oddflag bits), then project_N.  With
 * npre > 0 the last npre segments (the ray's first steps from the camera)
 * are read from pre_map, pre_flag instead. */
IPG_HD int chain_ray_segmented(const ipgpu_geom &G, const ipgpu_image &P, int ns, int seg_len,
                               const double *segmap, const int *segflag, double js, double ks, double out7[7],
                               int &nrad, const double *pre_map = nullptr, const int *pre_flag = nullptr,
                               int npre = 0)
{
  SegChain c;
  for (int q = 0; q < 4; q++) c.N.d[q] = 0.;
  for (int q = 0; q < 6; q++) c.N.re[q] = c.N.im[q] = 0.;
  c.I = c.tau = c.tauF = 0.;
  c.js = js;
  c.ks = ks;
  c.flag = c.nrad = 0;
  const int nseg = (ns + seg_len - 1) / seg_len;
  const int ntail = nseg - npre;
  for (int sgi = 0; sgi < ntail; sgi++) compose_segment(c, segmap + (size_t)sgi * SEGCOL * NPAR, segflag[sgi]);
  for (int sgi = 0; sgi < npre; sgi++) compose_segment(c, pre_map + (size_t)sgi * SEGCOL * NPAR, pre_flag[sgi]);
  Herm &N = c.N;
  int flag = c.flag;
  const double I = c.I, tau = c.tau, tauF = c.tauF;
  nrad = c.nrad;
  if (tauF > 1.e100 || tauF < -1.e100 || isnan(tauF)) flag |= 2;
  if (isnan(N.d[0])) flag |= 4;
  double S[4];
  double E[4][4];
  Metric m;
  const double Xc[4] = {P.Xcam[0], P.Xcam[1], P.Xcam[2], P.Xcam[3]};
  camera_tetrad(G, Xc, 0, E, m);
  project(N, E, m, P.rotcam, S);
  out7[0] = I;
  out7[1] = S[0];
  out7[2] = S[1];
  out7[3] = S[2];
  out7[4] = S[3];
  out7[5] = tau;
  out7[6] = tauF;
  return flag;
}

/* The fluid state at a fused transfer step: the mixed-precision instance
 * (F32) with d32 given gathers the FP32 copy of the cell records, any other
 * the FP64 records. */
template <bool F32>
IPG_HD_INLINE void fluid_step(const ipgpu_geom &G, const ipgpu_model &Mo, const double *dump, const float *d32,
                              const double X[4], const Metric &m, Fluid &F)
{
  double v[10];
  bool pre = false;
  if (F32 && d32 && Mo.kind != IPGPU_MODEL_SPHERE && in_domain(G, X[1], X[2])) {
    Cell c;
    locate(G, X[1], X[2], X[3], c);
    gather10_f32(G, d32, c, v);
    pre = true;
  }
  fluid(G, Mo, dump, X, m, F, nullptr, v, pre);
}

/* One transfer step fused: the work of prep_step and chain_step in the
 * order that keeps the fewest values live, with nothing written to memory.
 * The same operations on the same values as the split path; the half-step
 * metric and the end-point transport matrix are only formed while N is
 * nonzero (transport of N = 0 is the identity). */
template <bool F32 = false>
IPG_HD_INLINE void fused_step(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P, int n,
                              const double *rec, const int *nt, const double *dump, const Scale &sc,
                              bool pol, Chain &st, double (*Mi)[4], const float *d32)
{
  Rec ti, tf;
  load_rec(rec, n, ti);
  load_rec(rec, n - 1, tf);
  scale_rec(ti, sc);
  scale_rec(tf, sc);
  const double conv = Mo.L_unit * HPL / (ME * CL * CL);
  const double dl_cgs = ti.dl * conv;
  Metric m;
  metric(G, tf.X[1], tf.X[2], m);
  if (pol && st.Nnz) {
    /* push_polar twice, src/ipolarray.c:85-89 */
    double Mh[4][4];
    {
      double Xh[4], Kh[4];
      half_step(m, tf.X, tf.K, -ti.dl, Xh, Kh);
      Metric mh;
      metric(G, Xh[1], Xh[2], mh);
      transport_matrix(mh, Kh, Mh);
    }
    Herm Nh;
    transport(st.N, st.N, Mi, (0.5 * dl_cgs) / conv, Nh);
    transport(st.N, Nh, Mh, dl_cgs / conv, st.N);
  }
  /* FP64: the next step's transport matrix at this step's end point, formed
   * here so that the metric's derivatives are not live through the rest of
   * the step (transport only reads it while N != 0, so forming it for every
   * step gives the same values); the mixed variant has the registers and
   * forms it at the end, while N != 0 only */
  if (!F32 && pol) transport_matrix(m, tf.K, Mi);
  if (radiating_region(G, Mo, m.r, m.th)) {
    st.nrad++;
    bool zero_emission = false;
    if (P.target_nturns >= 0 && nt && nt[n] != P.target_nturns) zero_emission = true;
    const double half_x2 = (G.cstopx[2] - G.cstartx[2]) / 2;
    if (P.counterjet == 1) {
      if (tf.X[2] < half_x2) zero_emission = true;
    } else if (P.counterjet == 2) {
      if (tf.X[2] > half_x2) zero_emission = true;
    }
    Fluid F;
    fluid_step<F32>(G, Mo, dump, d32, tf.X, m, F);
    F.ne *= sc.ne;
    F.B *= sc.b;
    Coef c;
    double jf, kf;
    int flag = 0;
    double S0[4], E1[4], E2[4];
    if constexpr (F32) {
      coefficients_f32(F, tf.K, P.emission_type, pol, jf, kf, c);
    } else if (pol) {
      /* evolve_N's plasma frame (src/ipolarray.c:370-388) before the fits:
       * the four-vectors, the metric and the lowered tetrad rows are then
       * dead while the fits and the Stokes map are formed */
      const Geo q = F.ne > 0. ? coef_geometry(F, tf.K) : Geo{0., 1., 0., false};
      double Bcon[4] = {F.Bcon[0], F.Bcon[1], F.Bcon[2], F.Bcon[3]};
      double bsq = F.Bcon[0] * F.Bcov[0] + F.Bcon[1] * F.Bcov[1] + F.Bcon[2] * F.Bcov[2]
          + F.Bcon[3] * F.Bcov[3];
      if (bsq <= 0.) {
        Bcon[0] = 0.;
        Bcon[1] = 1.;
        Bcon[2] = 1.;
        Bcon[3] = 1.;
      }
      double E[4][4];
      flag = plasma_tetrad(F.Ucon, tf.K, Bcon, m.g, m.sqrtmg, E);
      if (isnan(F.Ucov[0])) flag |= 8;
      double e1[4], e2[4];
      lower(m.g, E[1], e1);
      lower(m.g, E[2], e2);
      herm_to_stokes(st.N, e1, e2, S0);
      for (int k = 0; k < 4; k++) {
        E1[k] = E[1][k];
        E2[k] = E[2][k];
      }
      if (F.ne > 0.) coefficients_core(q, F.ne, F.thetae, F.B, P.emission_type, pol, jf, kf, c);
      else coefficients_core(Geo{0., 1., 0., false}, 0., 1., 0., P.emission_type, pol, jf, kf, c);
    } else {
      coefficients(F, tf.K, P.emission_type, pol, jf, kf, c);
    }
    double ji = zero_emission ? 0. : st.js, ki = st.ks;
    st.js = jf;
    st.ks = kf;
    st.I = approximate_solve(st.I, ji, ki, zero_emission ? 0. : jf, kf, dl_cgs, st.tau);
    if (pol) {
      if (zero_emission) c.jI = c.jQ = c.jU = c.jV = 0.;
      if constexpr (F32) {
        /* evolve_N, src/ipolarray.c:332-575 */
        double Bcon[4] = {F.Bcon[0], F.Bcon[1], F.Bcon[2], F.Bcon[3]};
        double bsq = F.Bcon[0] * F.Bcov[0] + F.Bcon[1] * F.Bcov[1] + F.Bcon[2] * F.Bcov[2]
            + F.Bcon[3] * F.Bcov[3];
        if (bsq <= 0.) {
          Bcon[0] = 0.;
          Bcon[1] = 1.;
          Bcon[2] = 1.;
          Bcon[3] = 1.;
        }
        double E[4][4];
        flag = plasma_tetrad(F.Ucon, tf.K, Bcon, m.g, m.sqrtmg, E);
        if (isnan(F.Ucov[0])) flag |= 8;
        double e1[4], e2[4];
        lower(m.g, E[1], e1);
        lower(m.g, E[2], e2);
        herm_to_stokes(st.N, e1, e2, S0);
        for (int k = 0; k < 4; k++) {
          E1[k] = E[1][k];
          E2[k] = E[2][k];
        }
      }
      double T[4][4], cv[4];
      if (F32) stokes_affine_f32(c, dl_cgs, T, cv);
      else stokes_affine(c, dl_cgs, T, cv);
      double S[4];
      for (int i = 0; i < 4; i++) S[i] = cv[i] + T[i][0] * S0[0] + T[i][1] * S0[1] + T[i][2] * S0[2] + T[i][3] * S0[3];
      st.tauF += dl_cgs * fabs(c.rV);
      if (P.stokes_floors) {
        if (S[0] < 0) {
          S[0] = S[1] = S[2] = S[3] = 0.;
        } else {
          double pf = sqrt(S[1] * S[1] + S[2] * S[2] + S[3] * S[3]) / S[0];
          if (pf > 1.) {
            S[1] /= pf;
            S[2] /= pf;
            S[3] /= pf;
          }
        }
      }
      double dS[4] = {S[0] - S0[0], S[1] - S0[1], S[2] - S0[2], S[3] - S0[3]};
      herm_add_stokes(st.N, E1, E2, dS);
      if (st.tauF > 1.e100 || st.tauF < -1.e100 || isnan(st.tauF)) flag |= 2;
      if (isnan(st.N.d[0]) || isnan(S[0])) flag |= 4;
      st.flag |= flag;
      st.Nnz = herm_nonzero(st.N);
    }
  }
  if (F32 && pol && st.Nnz) transport_matrix(m, tf.K, Mi);
}

/* Whole transfer of one ray, step by step in registers (no packet memory):
 * for launches with enough rays to fill the chip.  mi (may be null: the
 * state's own) holds the 16 doubles of the transport matrix carried from
 * one step to the next; a kernel passes a slot of LDS there to free the
 * registers. */
template <bool F32 = false>
IPG_HD_INLINE int transfer_fused(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P, int ns,
                          const double *rec, const int *nt, const double *dump, const Scale &sc,
                          double out7[7], int &nrad, double *mi = nullptr, const float *d32 = nullptr)
{
  const bool pol = !P.only_unpolarized;
  double js, ks;
  far_end_coefficients<F32>(G, Mo, P, ns, rec, dump, js, ks, sc);
  Chain st;
  chain_init(st, js, ks);
  double(*Mi)[4] = mi ? reinterpret_cast<double(*)[4]>(mi) : st.Mi;
  for (int q = 0; q < ns; q++) fused_step<F32>(G, Mo, P, ns - q, rec, nt, dump, sc, pol, st, Mi, d32);
  nrad = st.nrad;
  return chain_finish(G, P, st, out7);
}

/* ------------------------------------------------ several jobs of one ray
 *
 * The jobs of a batch that share a geodesic trace (same camera, different
 * frequencies, M_unit values or image options) share every part of a
 * transfer step that depends on the geometry alone: the metric at the step's
 * end and half-step points, the transport matrices (Gamma K is linear in K,
 * and its product with the step dl, the only way it is used, is invariant
 * under K -> f K, dl -> dl / f), the fluid state at the point, the
 * fluid-frame angle and the plasma tetrad (functions of K's direction).
 * transfer_fused_multi runs up to MJ jobs of one ray in one lane: the shared
 * part once per step, then per job the fits at its frequency (nu scaled by
 * f), n_e and B scales, the Stokes step with its dl, and its coherency tensor,
 * which the caller keeps in memory (nst: element e of job j at
 * nst[(16 j + e) * stride]; on the device a lane's slots in LDS).  Each job's
 * results agree with transfer_fused on the job's rescaled records to the
 * roundoff of the rescaling. */
constexpr int MJ = 3;
/* job loops unrolled, so that the per-job arrays are indexed by constants
 * (a run-time index would put them in scratch memory) */
#if defined(__HIP_DEVICE_COMPILE__)
#define IPG_UNROLL_MJ _Pragma("unroll")
#else
#define IPG_UNROLL_MJ
#endif

struct JobCfg {
  Scale sc;
  int etype, pol, floors, cj;
};

struct JobAcc {
  double I, tau, tauF, js, ks;
  int flag, nrad, Nnz;
};

IPG_HD_INLINE void nst_load(const double *nst, int stride, int j, Herm &N)
{
#pragma unroll
  for (int e = 0; e < 16; e++) herm_par(N, e) = nst[(size_t)(16 * j + e) * stride];
}

IPG_HD_INLINE void nst_store(double *nst, int stride, int j, const Herm &N)
{
#pragma unroll
  for (int e = 0; e < 16; e++) nst[(size_t)(16 * j + e) * stride] = herm_par(N, e);
}

template <bool F32 = false>
IPG_HD_INLINE void fused_step_multi(const ipgpu_geom &G, const ipgpu_model &Mo, int nj, const JobCfg *cfg,
                                    int target_nturns, int n, const Rec &ti, const Rec &tf, const int *nt,
                                    const double *dump, JobAcc *acc, double *nst, int stride, double (*Mi)[4],
                                    bool any_pol, const float *d32)
{
  const double conv = Mo.L_unit * HPL / (ME * CL * CL);
  Metric m;
  metric(G, tf.X[1], tf.X[2], m);
  bool anyN = false;
  IPG_UNROLL_MJ
  for (int j = 0; j < MJ; j++) anyN = anyN || (j < nj && cfg[j].pol && acc[j].Nnz);
  if (anyN) {
    /* push_polar twice, src/ipolarray.c:85-89, with the trace's K and dl */
    double Mh[4][4];
    {
      double Xh[4], Kh[4];
      half_step(m, tf.X, tf.K, -ti.dl, Xh, Kh);
      Metric mh;
      metric(G, Xh[1], Xh[2], mh);
      transport_matrix(mh, Kh, Mh);
    }
    const double dl_cgs = ti.dl * conv;
    IPG_UNROLL_MJ
    for (int j = 0; j < MJ; j++) {
      if (j >= nj || !(cfg[j].pol && acc[j].Nnz)) continue;
      Herm N, Nh;
      nst_load(nst, stride, j, N);
      transport(N, N, Mi, (0.5 * dl_cgs) / conv, Nh);
      transport(N, Nh, Mh, dl_cgs / conv, N);
      nst_store(nst, stride, j, N);
    }
  }
  if (!F32 && any_pol) transport_matrix(m, tf.K, Mi);
  bool anyNnz = false;
  if (radiating_region(G, Mo, m.r, m.th)) {
    Fluid F;
    fluid_step<F32>(G, Mo, dump, d32, tf.X, m, F);
    const Geo q0 = F.ne > 0. ? coef_geometry(F, tf.K) : Geo{0., 1., 0., false};
    double E1[4], E2[4], e1[4], e2[4];
    int tflag = 0;
    if (any_pol) {
      /* evolve_N's plasma frame (src/ipolarray.c:370-388) */
      double Bcon[4] = {F.Bcon[0], F.Bcon[1], F.Bcon[2], F.Bcon[3]};
      double bsq = F.Bcon[0] * F.Bcov[0] + F.Bcon[1] * F.Bcov[1] + F.Bcon[2] * F.Bcov[2] + F.Bcon[3] * F.Bcov[3];
      if (bsq <= 0.) {
        Bcon[0] = 0.;
        Bcon[1] = 1.;
        Bcon[2] = 1.;
        Bcon[3] = 1.;
      }
      double E[4][4];
      tflag = plasma_tetrad(F.Ucon, tf.K, Bcon, m.g, m.sqrtmg, E);
      if (isnan(F.Ucov[0])) tflag |= 8;
      lower(m.g, E[1], e1);
      lower(m.g, E[2], e2);
      for (int k = 0; k < 4; k++) {
        E1[k] = E[1][k];
        E2[k] = E[2][k];
      }
    }
    const double half_x2 = (G.cstopx[2] - G.cstartx[2]) / 2;
    bool zero_turns = target_nturns >= 0 && nt && nt[n] != target_nturns;
    IPG_UNROLL_MJ
    for (int j = 0; j < MJ; j++) {
      if (j >= nj) break;
      const JobCfg &J = cfg[j];
      JobAcc &a = acc[j];
      a.nrad++;
      bool zero_emission = zero_turns;
      if (J.cj == 1) {
        if (tf.X[2] < half_x2) zero_emission = true;
      } else if (J.cj == 2) {
        if (tf.X[2] > half_x2) zero_emission = true;
      }
      const double dl_cgs = (ti.dl * J.sc.ikf) * conv;
      const double ne = F.ne * J.sc.ne, B = F.B * J.sc.b;
      Geo q = q0;
      q.nu *= J.sc.kf;
      Coef c;
      double jf, kf;
      if (ne > 0.) {
        if (F32) coefficients_f32_core(q, ne, F.thetae, B, J.etype, J.pol, jf, kf, c);
        else coefficients_core(q, ne, F.thetae, B, J.etype, J.pol, jf, kf, c);
      } else {
        if (F32) coefficients_f32_core(Geo{0., 1., 0., false}, 0., 1., 0., J.etype, J.pol, jf, kf, c);
        else coefficients_core(Geo{0., 1., 0., false}, 0., 1., 0., J.etype, J.pol, jf, kf, c);
      }
      const double ji = zero_emission ? 0. : a.js, ki = a.ks;
      a.js = jf;
      a.ks = kf;
      a.I = approximate_solve(a.I, ji, ki, zero_emission ? 0. : jf, kf, dl_cgs, a.tau);
      if (J.pol) {
        if (zero_emission) c.jI = c.jQ = c.jU = c.jV = 0.;
        Herm N;
        nst_load(nst, stride, j, N);
        double S0[4];
        herm_to_stokes(N, e1, e2, S0);
        double T[4][4], cv[4];
        if (F32) stokes_affine_f32(c, dl_cgs, T, cv);
        else stokes_affine(c, dl_cgs, T, cv);
        double S[4];
        for (int i = 0; i < 4; i++)
          S[i] = cv[i] + T[i][0] * S0[0] + T[i][1] * S0[1] + T[i][2] * S0[2] + T[i][3] * S0[3];
        a.tauF += dl_cgs * fabs(c.rV);
        if (J.floors) {
          if (S[0] < 0) {
            S[0] = S[1] = S[2] = S[3] = 0.;
          } else {
            double pf = sqrt(S[1] * S[1] + S[2] * S[2] + S[3] * S[3]) / S[0];
            if (pf > 1.) {
              S[1] /= pf;
              S[2] /= pf;
              S[3] /= pf;
            }
          }
        }
        double dS[4] = {S[0] - S0[0], S[1] - S0[1], S[2] - S0[2], S[3] - S0[3]};
        herm_add_stokes(N, E1, E2, dS);
        int flag = tflag;
        if (a.tauF > 1.e100 || a.tauF < -1.e100 || isnan(a.tauF)) flag |= 2;
        if (isnan(N.d[0]) || isnan(S[0])) flag |= 4;
        a.flag |= flag;
        a.Nnz = herm_nonzero(N);
        nst_store(nst, stride, j, N);
      }
      anyNnz = anyNnz || (J.pol && a.Nnz);
    }
  } else {
    IPG_UNROLL_MJ
    for (int j = 0; j < MJ; j++) anyNnz = anyNnz || (j < nj && cfg[j].pol && acc[j].Nnz);
  }
  if (F32 && anyNnz) transport_matrix(m, tf.K, Mi);
}

/* The whole transfer of one ray for nj <= MJ jobs (cfg, their images P[j]
 * for the projection), outputs out7[7 j ..]; returns nothing, flags and
 * radiating-step counts per job in flags[j], nrad[j].  mi: the transport
 * matrix carried between steps (16 doubles), nst: the jobs' coherency
 * tensors (see above). */
template <bool F32 = false>
IPG_HD_INLINE void transfer_fused_multi(const ipgpu_geom &G, const ipgpu_model &Mo, int nj, const JobCfg *cfg,
                                        const ipgpu_image *const *P, int ns, const double *rec, const int *nt,
                                        const double *dump, double *out7, int *flags, int *nrad, double *mi,
                                        double *nst, int stride, const float *d32 = nullptr)
{
  JobAcc acc[MJ];
  bool any_pol = false;
  {
    /* the far end's unpolarized coefficients, per job */
    Rec t;
    load_rec(rec, ns, t);
    Metric m;
    metric(G, t.X[1], t.X[2], m);
    Fluid F;
    fluid(G, Mo, dump, t.X, m, F);
    const Geo q0 = F.ne > 0. ? coef_geometry(F, t.K) : Geo{0., 1., 0., false};
    IPG_UNROLL_MJ
    for (int j = 0; j < MJ; j++) {
      if (j >= nj) break;
      const JobCfg &J = cfg[j];
      const double ne = F.ne * J.sc.ne, B = F.B * J.sc.b;
      Geo q = q0;
      q.nu *= J.sc.kf;
      Coef c;
      double js, ks;
      if (ne > 0.) {
        if (F32) coefficients_f32_core(q, ne, F.thetae, B, J.etype, false, js, ks, c);
        else coefficients_core(q, ne, F.thetae, B, J.etype, false, js, ks, c);
      } else {
        if (F32) coefficients_f32_core(Geo{0., 1., 0., false}, 0., 1., 0., J.etype, false, js, ks, c);
        else coefficients_core(Geo{0., 1., 0., false}, 0., 1., 0., J.etype, false, js, ks, c);
      }
      JobAcc &a = acc[j];
      a.I = a.tau = a.tauF = 0.;
      a.js = js;
      a.ks = ks;
      a.flag = a.nrad = a.Nnz = 0;
      Herm N;
      for (int q2 = 0; q2 < 4; q2++) N.d[q2] = 0.;
      for (int q2 = 0; q2 < 6; q2++) N.re[q2] = N.im[q2] = 0.;
      nst_store(nst, stride, j, N);
      any_pol = any_pol || J.pol;
    }
  }
  double(*Mi)[4] = reinterpret_cast<double(*)[4]>(mi);
  const int tn = P[0]->target_nturns;
  /* a step's end record is loaded one step ahead (the records of step n:
   * n at its start, n - 1 at its end), so that its latency hides behind the
   * step before (mixed transfer -2.5%); index 0 stands in past the last
   * step.  (In transfer_fused the carried record makes the mixed-precision
   * kernels spill, so it loads its records in the step.) */
  Rec ti, tf;
  load_rec(rec, ns, ti);
  load_rec(rec, ns >= 1 ? ns - 1 : 0, tf);
  for (int q = 0; q < ns; q++) {
    const int n = ns - q;
    Rec ahead;
    load_rec(rec, n >= 2 ? n - 2 : 0, ahead);
    fused_step_multi<F32>(G, Mo, nj, cfg, tn, n, ti, tf, nt, dump, acc, nst, stride, Mi, any_pol, d32);
    ti = tf;
    tf = ahead;
  }
  IPG_UNROLL_MJ
  for (int j = 0; j < MJ; j++) {
    if (j >= nj) break;
    Chain st;
    nst_load(nst, stride, j, st.N);
    st.I = acc[j].I;
    st.tau = acc[j].tau;
    st.tauF = acc[j].tauF;
    st.flag = acc[j].flag;
    flags[j] = chain_finish(G, *P[j], st, out7 + 7 * j);
    nrad[j] = acc[j].nrad;
  }
}

#if !defined(IPG_OPCOUNT)
/* ------------------------------------------------------------------ adjoint
 *
 * Reverse mode through the state recurrence of one ray, for the gradient of
 * L = sum over pixels and I, Q, U, V of w_s (S_s - target_s)^2 (S in the
 * reference's output units and sign convention) with respect to M_unit,
 * R_high, R_low and a per-cell scale of n_e.  The adjoint state is a
 * Hermitian Lambda with dL = Re tr(Lambda^dagger dN); the transposes of the
 * step's linear maps are:
 *   transport  N -> N - dl (M N + N M^T):        Lambda -> Lambda - dl (M^T Lambda + Lambda M)
 *   Stokes     S = K_e(N) (herm_to_stokes, rows e1, e2):  K_e^T(s) = herm_add_stokes(e1, e2, s / 2)
 *   update     N += L_E(dS) (herm_add_stokes, E1, E2):    L_E^T(Lambda) = 2 herm_to_stokes(Lambda, E1, E2)
 * The step's dependence on n_e, Theta_e and B (through the fits and the
 * exact Stokes step) is taken by central differences of the coefficient
 * map at the step (relative step 1e-5).  States are checkpointed every CK
 * steps on the forward pass and recomputed segment by segment.  The Stokes
 * floors, a non-differentiable clamp that is off by default, are not
 * differentiated. */
constexpr int CK = 32;

/* M N + N M^T (no scaling) */
IPG_HD void herm_LM(const Herm &N, const double M[4][4], bool transposed, Herm &out)
{
  double Mt[4][4];
  for (int i = 0; i < 4; i++)
    for (int k = 0; k < 4; k++) Mt[i][k] = transposed ? M[k][i] : M[i][k];
  Herm zero;
  for (int q = 0; q < 4; q++) zero.d[q] = 0.;
  for (int q = 0; q < 6; q++) zero.re[q] = zero.im[q] = 0.;
  transport(zero, N, Mt, -1., out);
}

IPG_HD void herm_axpy(Herm &y, double a, const Herm &x)
{
  for (int q = 0; q < 4; q++) y.d[q] += a * x.d[q];
  for (int q = 0; q < 6; q++) {
    y.re[q] += a * x.re[q];
    y.im[q] += a * x.im[q];
  }
}

IPG_HD void atomic_add_f64(double *p, double v)
{
#if defined(__HIP_DEVICE_COMPILE__)
  atomicAdd(p, v);
#else
#pragma omp atomic
  *p += v;
#endif
}

/* The reverse pass over steps [q0, q1) of one ray's recurrence: from the
 * state Nq0 before step q0 (checkpointed every CK steps on a forward pass,
 * recomputed window by window) and the adjoint state L after step q1 - 1,
 * accumulates g3 += (dL/dln M_unit, dL/dR_high, dL/dR_low) and the per-cell
 * gradients, and leaves in L the adjoint state before step q0.  pkt/aux:
 * the ray's packets and adjoint records (packet q is step ns - q); ckpt:
 * room for ceil((q1 - q0) / CK) states; seg: room for CK states. */
IPG_HD void adjoint_range(const ipgpu_geom &G, const ipgpu_image &P, const double *pkt, const double *aux, int q0,
                          int q1, const Herm &Nq0, Herm &L, const double *dump, Herm *ckpt, Herm *seg, double *gcell,
                          double g3[3])
{
  Chain st;
  chain_init(st, 0., 0.);
  st.N = Nq0;
  st.Nnz = herm_nonzero(Nq0);
  if (q0 > 0)
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) st.Mi[i][j] = pkt[(size_t)(q0 - 1) * PK + PK_MF + 4 * i + j];
  for (int q = q0; q < q1; q++) {
    if ((q - q0) % CK == 0) ckpt[(q - q0) / CK] = st.N;
    chain_step(st, pkt + (size_t)q * PK, true, 0);
  }
  const long s2 = G.N3 + 2, s1 = (long)(G.N2 + 2) * (G.N3 + 2);
  const int nwin = (q1 - q0 + CK - 1) / CK;
  for (int k = nwin - 1; k >= 0; k--) {
    const int w0 = q0 + k * CK, w1 = (w0 + CK < q1) ? w0 + CK : q1;
    Chain rs;
    chain_init(rs, 0., 0.);
    rs.N = ckpt[k];
    rs.Nnz = herm_nonzero(rs.N);
    if (w0 > 0)
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) rs.Mi[i][j] = pkt[(size_t)(w0 - 1) * PK + PK_MF + 4 * i + j];
    for (int q = w0; q < w1; q++) {
      seg[q - w0] = rs.N;
      chain_step(rs, pkt + (size_t)q * PK, true, 0);
    }
    for (int q = w1 - 1; q >= w0; q--) {
      const double *pk = pkt + (size_t)q * PK;
      const double *ax = aux + (size_t)q * AX;
      const Herm &Nb = seg[q - w0];
      const bool nnz = herm_nonzero(Nb);
      double Mi[4][4], Mh[4][4];
      if (nnz) {
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++) {
            Mi[i][j] = pkt[(size_t)(q - 1) * PK + PK_MF + 4 * i + j];
            Mh[i][j] = pk[PK_MH + 4 * i + j];
          }
      }
      const int meta = (int)pk[PK_META];
      if (meta & 1) {
        Herm Nt = Nb;
        if (nnz) {
          Herm Nh;
          transport(Nb, Nb, Mi, pk[PK_HDL], Nh);
          transport(Nb, Nh, Mh, pk[PK_FDL], Nt);
        }
        double mu[4], S0[4];
        herm_to_stokes(L, pk + PK_E1, pk + PK_E2, mu);
        for (int i = 0; i < 4; i++) mu[i] *= 2.;
        herm_to_stokes(Nt, pk + PK_e1, pk + PK_e2, S0);
        if (ax[AX_NE] > 0.) {
          /* g_p = mu . (dT/dp S0 + dcv/dp) with the maps the prep pass formed */
          double g[3];
          for (int p = 0; p < 3; p++) {
            const double *dT = ax + AX_DT + 16 * p, *dc = ax + AX_DCV + 4 * p;
            double acc = 0.;
            for (int i = 0; i < 4; i++)
              acc += mu[i] * (dc[i] + dT[4 * i] * S0[0] + dT[4 * i + 1] * S0[1] + dT[4 * i + 2] * S0[2]
                              + dT[4 * i + 3] * S0[3]);
            g[p] = acc;
          }
          g3[0] += g[0] + 0.5 * g[2];
          g3[1] += g[1] * ax[AX_DRH];
          g3[2] += g[1] * ax[AX_DRL];
          if (gcell && ax[AX_CELL] >= 0.) {
            const long base = (long)ax[AX_CELL];
            const double d[3] = {ax[AX_D1], ax[AX_D2], ax[AX_D3]};
            const double f = g[0] * ax[AX_SMOOTH] / ax[AX_NE];
            for (int c = 0; c < 8; c++) {
              const int di = (c >> 2) & 1, dj = (c >> 1) & 1, dk = c & 1;
              const long idx = base + di * s1 + dj * s2 + dk;
              const double wgt = (di ? d[0] : 1. - d[0]) * (dj ? d[1] : 1. - d[1]) * (dk ? d[2] : 1. - d[2]);
              const double ne_c = dump[CELL_STRIDE * idx + 6];
              if (wgt * ne_c != 0.) atomic_add_f64(gcell + idx, f * wgt * ne_c);
            }
          }
        }
        /* Lambda += K_e^T((T - I)^T mu) */
        double v[4];
        for (int i = 0; i < 4; i++) {
          double acc = -mu[i];
          for (int r = 0; r < 4; r++) acc += pk[PK_T + 4 * r + i] * mu[r];
          v[i] = 0.5 * acc;
        }
        herm_add_stokes(L, pk + PK_e1, pk + PK_e2, v);
      }
      if (nnz) {
        /* transport^T: Lambda - f G + f h L_{Mi^T}(G), G = L_{Mh^T}(Lambda) */
        Herm Gm, Hm;
        herm_LM(L, Mh, true, Gm);
        herm_LM(Gm, Mi, true, Hm);
        const double f = pk[PK_FDL], h = pk[PK_HDL];
        herm_axpy(L, -f, Gm);
        herm_axpy(L, f * h, Hm);
      }
    }
  }
}

/* The pixel's I, Q, U, V (output units) from the final state N and the
 * adjoint state the loss puts on N: with cot null, the L2 loss against tgt
 * with weights w (loss returned), else the given dL/d(I, Q, U, V). */
IPG_HD double adjoint_seed(const ipgpu_geom &G, const ipgpu_image &P, const Herm &N, const double tgt[4],
                           const double w[4], const double *cot, double Simg[4], Herm &L)
{
  double E[4][4];
  Metric mc;
  const double Xc[4] = {P.Xcam[0], P.Xcam[1], P.Xcam[2], P.Xcam[3]};
  camera_tetrad(G, Xc, 0, E, mc);
  double Sc[4];
  project(N, E, mc, P.rotcam, Sc);
  const double f3 = P.freqcgs * P.freqcgs * P.freqcgs;
  const double qs = (P.qu_conv == 0) ? -1. : 1.;
  const double sgn[4] = {1., qs, qs, 1.};
  double sig[4], loss = 0.;
  for (int i = 0; i < 4; i++) {
    Simg[i] = f3 * sgn[i] * Sc[i];
    if (cot) {
      sig[i] = cot[i] * f3 * sgn[i];
    } else {
      double r = Simg[i] - tgt[i];
      loss += w[i] * r * r;
      sig[i] = 2. * w[i] * r * f3 * sgn[i];
    }
  }
  /* project_N^T: the rotation by -2 rotcam, then K_e^T with the camera rows */
  const double rc = -2. * P.rotcam, cr = cos(rc), sr = sin(rc);
  double tau[4] = {0.5 * sig[0], 0.5 * (sig[1] * cr + sig[2] * sr), 0.5 * (-sig[1] * sr + sig[2] * cr), 0.5 * sig[3]};
  double e1c[4], e2c[4];
  lower(mc.g, E[1], e1c);
  lower(mc.g, E[2], e2c);
  for (int q = 0; q < 4; q++) L.d[q] = 0.;
  for (int q = 0; q < 6; q++) L.re[q] = L.im[q] = 0.;
  herm_add_stokes(L, e1c, e2c, tau);
  return loss;
}

/* Loss and gradient of one ray in one lane (sequential: forward pass,
 * then the reverse pass over all steps).  pkt/aux: its packets and adjoint
 * records (packet 0 is step ns); tgt, w, cot: as adjoint_seed; ckpt: room
 * for ceil(ns / CK) states; seg: room for CK states; gcell (ghosted cells)
 * receives dL/d(n_e scale) of each cell.  out: loss, dL/dln(M_unit),
 * dL/dR_high, dL/dR_low; Simg: the pixel's I, Q, U, V. */
IPG_HD void adjoint_ray(const ipgpu_geom &G, const ipgpu_image &P, int ns, const double *pkt, const double *aux,
                        const double tgt[4], const double w[4], const double *cot, const double *dump, Herm *ckpt,
                        Herm *seg, double *gcell, double out[4], double Simg[4])
{
  Chain st;
  chain_init(st, 0., 0.);
  for (int q = 0; q < ns; q++) chain_step(st, pkt + (size_t)q * PK, true, 0);
  Herm L;
  out[0] = adjoint_seed(G, P, st.N, tgt, w, cot, Simg, L);
  Herm Z;
  for (int q = 0; q < 4; q++) Z.d[q] = 0.;
  for (int q = 0; q < 6; q++) Z.re[q] = Z.im[q] = 0.;
  double g3[3] = {0., 0., 0.};
  adjoint_range(G, P, pkt, aux, 0, ns, Z, L, dump, ckpt, seg, gcell, g3);
  out[1] = g3[0];
  out[2] = g3[1];
  out[3] = g3[2];
}

/* ---- the reverse pass in segments of a ray's steps (parallel in time)
 *
 * A ray's ns steps are cut into segments [k Ls, min((k+1) Ls, ns)) from the
 * far end.  The map of segment k, N -> A_k N + b_k, is formed column by
 * column (segment_column: A_k's 16 columns from the basis states, b_k from
 * N = 0 with sources; the first segment starts from N = 0 and needs b_0
 * only).  One lane per ray composes them: the state before every segment,
 * the final state and the loss's adjoint state, then the adjoint states
 * after every segment, backwards with the transposed maps.  The adjoint of
 * the linear map A in the parameters of Hermitian N (d[4], re[6], im[6]),
 * for the pairing Re tr(Lambda^dagger N) = sum_b W_b Lambda_b N_b with
 * W = 1 on the diagonal and 2 off it, is W^-1 A^T W.  Every segment's
 * reverse pass (adjoint_range) then runs in a lane of its own. */
constexpr int ASEG_MAP = (NPAR + 1) * NPAR;   /* doubles per segment map: 16 columns + source */

IPG_HD_INLINE double herm_w(int b) { return b < 4 ? 1. : 2.; }

/* out = A N + b with the map of one segment */
IPG_HD_INLINE void seg_map_apply(const double *map, const Herm &N, Herm &out)
{
#pragma unroll
  for (int c = 0; c < NPAR; c++) {
    double acc = map[NPAR * NPAR + c];
#pragma unroll
    for (int b = 0; b < NPAR; b++) acc += map[b * NPAR + c] * herm_par(N, b);
    herm_par(out, c) = acc;
  }
}

/* out = W^-1 A^T W L: the adjoint state before a segment from the one after */
IPG_HD_INLINE void seg_map_apply_T(const double *map, const Herm &L, Herm &out)
{
#pragma unroll
  for (int b = 0; b < NPAR; b++) {
    double acc = 0.;
#pragma unroll
    for (int c = 0; c < NPAR; c++) acc += map[b * NPAR + c] * herm_w(c) * herm_par(L, c);
    herm_par(out, b) = acc / herm_w(b);
  }
}

/* The composition of one ray of nseg segments (maps at map + k ASEG_MAP):
 * Nstart[k] the state before segment k, Lend[k] the adjoint state after
 * it; returns the loss (0 with cot) and the pixel's I, Q, U, V. */
IPG_HD double adjoint_compose(const ipgpu_geom &G, const ipgpu_image &P, int nseg, const double *map,
                              const double tgt[4], const double w[4], const double *cot, Herm *Nstart, Herm *Lend,
                              double Simg[4])
{
  Herm N;
  for (int q = 0; q < 4; q++) N.d[q] = 0.;
  for (int q = 0; q < 6; q++) N.re[q] = N.im[q] = 0.;
  for (int k = 0; k < nseg; k++) {
    Nstart[k] = N;
    if (k == 0) {
#pragma unroll
      for (int c = 0; c < NPAR; c++) herm_par(N, c) = map[NPAR * NPAR + c];
    } else {
      Herm t;
      seg_map_apply(map + (size_t)k * ASEG_MAP, N, t);
      N = t;
    }
  }
  Herm L;
  const double loss = adjoint_seed(G, P, N, tgt, w, cot, Simg, L);
  for (int k = nseg - 1; k >= 0; k--) {
    Lend[k] = L;
    if (k > 0) {
      Herm t;
      seg_map_apply_T(map + (size_t)k * ASEG_MAP, L, t);
      L = t;
    }
  }
  return loss;
}
#endif

/* ------------------------------------------------------- pixel trace output */

/* Everything dump_var_along records for one pixel (src/io.c:279-543), one
 * lane.  The geodesic is traced into rec (nstep + 1 records), then the
 * records are walked as the reference does: i = nstep - 1 .. 1, quantities
 * at point i, the coherency tensor before the step, then one transfer step
 * from point i + 1 to i with the emissivity cache started afresh from point
 * i + 1 (integrate_emission called with one step, src/io.c:428).  Fields per
 * record: the IPGPU_TR_* offsets of include/ipgpu.h.  rec holds cap
 * trajectory records, nt cap subring counts (read when target_nturns >= 0),
 * out cap zeroed output records; prims_ok says the dump carries rho, u and
 * beta (slots SLOT_RHO, SLOT_U, SLOT_BETA). */
IPG_HD void trace_pixel_records(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P, int pi,
                                int pj, int cap, int prims_ok, const double *dump, double *rec, int *nt,
                                double *out, int *nstep_out)
{
  const int ns = trace_ray(G, P, pi, pj, cap, rec, P.target_nturns >= 0 ? nt : nullptr);
  *nstep_out = ns;
  if (ns + 1 > cap) return;
  const int *ntp = P.target_nturns >= 0 ? nt : nullptr;
  const bool pol = !P.only_unpolarized;
  const double conv = Mo.L_unit * HPL / (ME * CL * CL);
  /* X0 along the ray: the time component the records leave out, formed by
   * the geodesic step's own arithmetic */
  {
    Rec r0;
    load_rec(rec, 0, r0);
    double X0 = P.Xcam[0];
    out[IPGPU_TR_X] = X0;
    for (int n = 1; n <= ns; n++) {
      Rec a, b;
      load_rec(rec, n - 1, a);
      load_rec(rec, n, b);
      double X[4] = {X0, a.X[1], a.X[2], a.X[3]};
      Metric m;
      metric(G, X[1], X[2], m);
      double Xh[4], Kh[4];
      half_step(m, X, a.K, -b.dl, Xh, Kh);
      X0 = X0 + (-b.dl) * Kh[0];
      out[(size_t)n * IPGPU_TRACE_NF + IPGPU_TR_X] = X0;
    }
  }
  Herm N;
  for (int q = 0; q < 4; q++) N.d[q] = 0.;
  for (int q = 0; q < 6; q++) N.re[q] = N.im[q] = 0.;
  for (int i = ns - 1; i > 0; --i) {
    double *o = out + (size_t)i * IPGPU_TRACE_NF;
    Rec t;
    load_rec(rec, i, t);
    double X[4] = {o[IPGPU_TR_X], t.X[1], t.X[2], t.X[3]};
    Metric m;
    metric(G, X[1], X[2], m);
    Fluid F;
    fluid(G, Mo, dump, X, m, F);
    /* primitives, sigma, beta */
    if (Mo.kind == IPGPU_MODEL_IHARM && in_domain(G, X[1], X[2])) {
      Cell c;
      locate(G, X[1], X[2], X[3], c);
      double v[10];
      gather10(G, dump, c, v);
      for (int q = 0; q < 6; q++) o[IPGPU_TR_PRIMS + 2 + q] = v[q];
      o[IPGPU_TR_SIGMA] = v[9];
      if (prims_ok) {
        o[IPGPU_TR_PRIMS + 0] = gather_slot(G, dump, c, SLOT_RHO);
        o[IPGPU_TR_PRIMS + 1] = gather_slot(G, dump, c, SLOT_U);
        o[IPGPU_TR_BETA] = gather_slot(G, dump, c, SLOT_BETA);
      }
    } else if (Mo.kind == IPGPU_MODEL_SPHERE && radiating_region(G, Mo, m.r, m.th)) {
      o[IPGPU_TR_SIGMA] = Mo.sph_sigma;
      o[IPGPU_TR_BETA] = Mo.sph_beta;
    }
    o[IPGPU_TR_B] = F.B;
    o[IPGPU_TR_NE] = F.ne;
    o[IPGPU_TR_THETAE] = F.thetae;
    const Geo q = coef_geometry(F, t.K);
    o[IPGPU_TR_NU] = q.nu;
    {
      /* get_bk_angle, src/radiation.c:47-72 */
      double Bm = sqrt(fabs(F.Bcon[0] * F.Bcov[0] + F.Bcon[1] * F.Bcov[1] + F.Bcon[2] * F.Bcov[2]
                            + F.Bcon[3] * F.Bcov[3]));
      double ang = PI / 2.;
      if (Bm != 0.) {
        double k = fabs(t.K[0] * F.Ucov[0] + t.K[1] * F.Ucov[1] + t.K[2] * F.Ucov[2] + t.K[3] * F.Ucov[3]);
        double mu = (t.K[0] * F.Bcov[0] + t.K[1] * F.Bcov[1] + t.K[2] * F.Bcov[2] + t.K[3] * F.Bcov[3]) / (k * Bm);
        if (fabs(mu) > 1.) mu /= fabs(mu);
        ang = acos(mu);
      }
      o[IPGPU_TR_BK_ANGLE] = ang;
    }
    /* jar_calc, get_jkinv */
    Coef c;
    double ju, ku;
    coefficients(F, t.K, P.emission_type, true, ju, ku, c);
    o[IPGPU_TR_J_INV + 0] = c.jI;
    o[IPGPU_TR_J_INV + 1] = c.jQ;
    o[IPGPU_TR_J_INV + 2] = c.jU;
    o[IPGPU_TR_J_INV + 3] = c.jV;
    o[IPGPU_TR_ALPHA_INV + 0] = c.aI;
    o[IPGPU_TR_ALPHA_INV + 1] = c.aQ;
    o[IPGPU_TR_ALPHA_INV + 2] = c.aU;
    o[IPGPU_TR_ALPHA_INV + 3] = c.aV;
    o[IPGPU_TR_RHO_INV + 1] = c.rQ;
    o[IPGPU_TR_RHO_INV + 2] = c.rU;
    o[IPGPU_TR_RHO_INV + 3] = c.rV;
    {
      Coef cu;
      coefficients(F, t.K, P.emission_type, false, ju, ku, cu);
      o[IPGPU_TR_J_UNPOL] = ju;
      o[IPGPU_TR_K_UNPOL] = ku;
    }
    o[IPGPU_TR_DL] = t.dl * conv;
    for (int k = 1; k < 4; k++) o[IPGPU_TR_X + k] = X[k];
    for (int k = 0; k < 4; k++) o[IPGPU_TR_KCON + k] = t.K[k];
    double rr, th;
    bl_coord(G, X[1], X[2], rr, th);
    o[IPGPU_TR_R] = rr;
    o[IPGPU_TR_TH] = th;
    o[IPGPU_TR_PHI] = X[3];
    /* plasma tetrad at point i (no b^2 <= 0 replacement here, as in the
     * reference's trace) */
    double E[4][4], Ec[4][4];
    plasma_tetrad(F.Ucon, t.K, F.Bcon, m.g, m.sqrtmg, E);
    for (int a = 0; a < 4; a++) lower(m.g, E[a], Ec[a]);
    for (int k = 0; k < 4; k++) Ec[0][k] = -Ec[0][k];
    for (int a = 0; a < 4; a++)
      for (int k = 0; k < 4; k++) {
        o[IPGPU_TR_ECON + 4 * a + k] = E[a][k];
        o[IPGPU_TR_ECOV + 4 * a + k] = Ec[a][k];
      }
    /* e2 carried by parallel_transport_vector twice (src/ipolarray.c:297-320,
     * src/io.c:393-408): to the step's half point with the connection at X_i,
     * then to X_{i-1} with the connection at the half point */
    {
      Rec tm;
      load_rec(rec, i - 1, tm);
      double Xp[4] = {out[(size_t)(i - 1) * IPGPU_TRACE_NF + IPGPU_TR_X], tm.X[1], tm.X[2], tm.X[3]};
      Metric mp;
      metric(G, Xp[1], Xp[2], mp);
      double Xh[4], Kh[4];
      half_step(mp, Xp, tm.K, -t.dl, Xh, Kh);   /* the geodesic step's half point (traj[i].Xhalf) */
      double M0[4][4], Mh[4][4];
      transport_matrix(m, t.K, M0);
      Metric mh;
      metric(G, Xh[1], Xh[2], mh);
      transport_matrix(mh, Kh, Mh);
      /* the reference passes the step in cgs and divides it back */
      const double dlc = (t.dl * conv) / conv;
      double e2m[4], e2f[4];
      for (int a = 0; a < 4; a++) {
        double sacc = 0.;
        for (int b = 0; b < 4; b++) sacc += M0[a][b] * E[2][b];
        e2m[a] = E[2][a] - sacc * dlc;
      }
      for (int a = 0; a < 4; a++) {
        double sacc = 0.;
        for (int b = 0; b < 4; b++) sacc += Mh[a][b] * e2m[b];
        e2f[a] = E[2][a] - sacc * dlc;
      }
      double *om = out + (size_t)(i - 1) * IPGPU_TRACE_NF;
      for (int k = 0; k < 4; k++) {
        o[IPGPU_TR_E2C + k] = E[2][k];
        o[IPGPU_TR_E2C_COV + k] = Ec[2][k];
        om[IPGPU_TR_E2T + k] = e2f[k];
      }
      /* e2Transported at i (carried from i + 1), lowered at X_i */
      double e2i[4] = {o[IPGPU_TR_E2T], o[IPGPU_TR_E2T + 1], o[IPGPU_TR_E2T + 2], o[IPGPU_TR_E2T + 3]};
      double e2c[4];
      lower(m.g, e2i, e2c);
      for (int k = 0; k < 4; k++) o[IPGPU_TR_E2T_COV + k] = e2c[k];
    }
    /* N before the step: coordinate and tetrad components (real parts), the
     * plasma-frame Stokes parameters */
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        double re, im;
        herm_get(N, a, b, re, im);
        o[IPGPU_TR_NCOORD + 4 * a + b] = re;
      }
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        double sacc = 0.;
        for (int mu = 0; mu < 4; mu++)
          for (int nu = 0; nu < 4; nu++) {
            double re, im;
            herm_get(N, mu, nu, re, im);
            sacc += Ec[a][mu] * Ec[b][nu] * re;
          }
        o[IPGPU_TR_NTETRAD + 4 * a + b] = sacc;
      }
    {
      double S[4];
      herm_to_stokes(N, Ec[1], Ec[2], S);
      for (int k = 0; k < 4; k++) o[IPGPU_TR_STOKES + k] = S[k];
    }
    /* integrate_emission over one step, from point i + 1 to i: I, tau and
     * tauF start from 0 and the cached coefficients from point i + 1; N
     * carries over */
    {
      Pkt pk;
      for (int f = 0; f < PK; f++) pk.v[f] = 0.;
      prep_step<false>(G, Mo, P, i + 1, rec, ntp, dump, pk.v, pol, unit_scale());
      Chain st;
      double js, ks;
      far_end_coefficients<false>(G, Mo, P, i + 1, rec, dump, js, ks, unit_scale());
      chain_init(st, js, ks);
      st.N = N;
      st.Nnz = herm_nonzero(N);
      Rec ti;
      load_rec(rec, i + 1, ti);
      Metric mi;
      metric(G, ti.X[1], ti.X[2], mi);
      transport_matrix(mi, ti.K, st.Mi);
      chain_step(st, pk.v, pol, P.stokes_floors);
      N = st.N;
      o[IPGPU_TR_I_UNPOL] = st.I;
    }
  }
  /* X holds index 1 .. nstep - 1 only */
  out[IPGPU_TR_X] = 0.;
  out[(size_t)ns * IPGPU_TRACE_NF + IPGPU_TR_X] = 0.;
}


}  // namespace ipg
