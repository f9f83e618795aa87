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
 * with a larger cap).  nt (may be null) receives the subring count per
 * step when P.target_nturns >= 0. */
IPG_HD int trace_ray(const ipgpu_geom &G, const ipgpu_image &P, int pi, int pj, int cap, double *rec,
                     int *nt)
{
  double E[4][4];
  Metric m;
  double X[4] = {P.Xcam[0], P.Xcam[1], P.Xcam[2], P.Xcam[3]};
  camera_tetrad(G, X, P.old_centering, E, m);
  double K[4];
  init_K(P, E, pi, pj, K);
  store_rec(rec, X, K, 0.);
  if (nt) nt[0] = 0;
  int passed = (P.Xcam[2] < 0.5) ? 1 : ((P.Xcam[2] > 0.5) ? -1 : 0);
  int nturns = 0;
  double za = 0., zb = 0., x2a = 0., x2b = 0.;
  int nstep = 0;
  while (true) {
    metric(G, X[1], X[2], m);
    if (stop_backward(G, m.r, K)) break;
    if (nstep >= P.maxnstep - 1) break;
    nstep++;
    double dl = stepsize(G, X, K, P.eps);
    double dlp = -dl;
    double Xh[4], Kh[4];
    half_step(m, X, K, dlp, Xh, Kh);
    metric(G, Xh[1], Xh[2], m);
    double acc[4];
    geodesic_accel(m, Kh, acc);
    for (int k = 0; k < 4; k++) K[k] = K[k] - dlp * acc[k];
    for (int k = 0; k < 4; k++) X[k] = X[k] + dlp * Kh[k];
    if (nstep < cap) store_rec(rec + (size_t)nstep * REC, X, K, dl);
    if (nt) {
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
  return nstep;
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

/* Packet of transfer step n (from trajectory point n to n-1) of
 * integrate_emission, src/ipolarray.c:78-180.  nt may be null. */
IPG_HD void prep_step(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P, int n,
                      const double *rec, const int *nt, const double *dump, double *pk, bool pol)
{
  Rec ti, tf;
  load_rec(rec, n, ti);
  load_rec(rec, n - 1, tf);
  const double conv = Mo.L_unit * HPL / (ME * CL * CL);
  const double dl_cgs = ti.dl * conv;
  Metric m;
  metric(G, tf.X[1], tf.X[2], m);
  int meta = 0;
  pk[PK_DL] = dl_cgs;
  if (pol) {
    double Mf[4][4], Mh[4][4], Xh[4], Kh[4];
    transport_matrix(m, tf.K, Mf);
    half_step(m, tf.X, tf.K, -ti.dl, Xh, Kh);
    Metric mh;
    metric(G, Xh[1], Xh[2], mh);
    transport_matrix(mh, Kh, Mh);
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 4; k++) {
        pk[PK_MF + 4 * i + k] = Mf[i][k];
        pk[PK_MH + 4 * i + k] = Mh[i][k];
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
    fluid(G, Mo, dump, tf.X, m, F);
    Coef c;
    double ju, ku;
    coefficients(F, tf.K, P.emission_type, pol, ju, ku, c);
    pk[PK_JU] = ju;
    pk[PK_KU] = ku;
    if (pol) {
      if (zero_emission) c.jI = c.jQ = c.jU = c.jV = 0.;
      double T[4][4], cv[4];
      stokes_affine(c, dl_cgs, T, cv);
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
IPG_HD void far_end_coefficients(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P,
                                 int ns, const double *rec, const double *dump, double &js, double &ks)
{
  Rec t;
  load_rec(rec, ns, t);
  Metric m;
  metric(G, t.X[1], t.X[2], m);
  Fluid F;
  fluid(G, Mo, dump, t.X, m, F);
  Coef c;
  coefficients(F, t.K, P.emission_type, false, js, ks, c);
}

/* One packet held in registers; on the device it is read with 16-byte
 * loads so that the next step's packet is in flight while the current step
 * computes. */
struct Pkt {
  double v[PK];
};

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

/* The state recurrence of integrate_emission over the packets of one ray
 * (packet 0 is step ns, packet ns-1 is step 1), then project_N.
 * out7 = Intensity, Is, Qs, Us, Vs, Tau, tauF (before the nu^3 scaling of
 * save_pixel); returns the oddflag bits; nrad counts radiating steps. */
IPG_HD int chain_ray(const ipgpu_geom &G, const ipgpu_image &P, int ns, const double *pkt, int stride,
                     double js, double ks, double out7[7], int &nrad)
{
  const bool pol = stride == PK;
  Herm N;
  for (int q = 0; q < 4; q++) N.d[q] = 0.;
  for (int q = 0; q < 6; q++) N.re[q] = N.im[q] = 0.;
  double I = 0., tau = 0., tauF = 0.;
  int flag = 0;
  nrad = 0;
  if (pol) {
    bool Nnz = false;
    double Mi[4][4];
    Pkt cur, nxt;
    if (ns > 0) load_pkt(pkt, PK, cur);
    for (int q = 0; q < ns; q++) {
      if (q + 1 < ns) load_pkt(pkt + (size_t)(q + 1) * PK, PK, nxt);
      const double *pk = cur.v;
      const int meta = (int)pk[PK_META];
      if (Nnz) {
        /* push_polar twice, src/ipolarray.c:85-89 */
        double Mh[4][4];
        for (int i = 0; i < 4; i++)
          for (int k = 0; k < 4; k++) Mh[i][k] = pk[PK_MH + 4 * i + k];
        Herm Nh;
        transport(N, N, Mi, pk[PK_HDL], Nh);
        transport(N, Nh, Mh, pk[PK_FDL], N);
      }
      if (meta & 1) {
        nrad++;
        double jf = pk[PK_JU], kf = pk[PK_KU];
        double ji = (meta & 2) ? 0. : js, ki = ks;
        js = jf;
        ks = kf;
        I = approximate_solve(I, ji, ki, (meta & 2) ? 0. : jf, kf, pk[PK_DL], tau);
        flag |= stokes_apply(N, pk, P.stokes_floors, tauF);
        flag |= meta >> 2;
        Nnz = herm_nonzero(N);
      }
      if (Nnz)
        for (int i = 0; i < 4; i++)
          for (int k = 0; k < 4; k++) Mi[i][k] = pk[PK_MF + 4 * i + k];
      cur = nxt;
    }
  } else {
    for (int q = 0; q < ns; q++) {
      const double *pk = pkt + (size_t)q * stride;
      const int meta = (int)pk[PK_META];
      if (meta & 1) {
        nrad++;
        double jf = pk[PK_JU], kf = pk[PK_KU];
        double ji = (meta & 2) ? 0. : js, ki = ks;
        js = jf;
        ks = kf;
        I = approximate_solve(I, ji, ki, (meta & 2) ? 0. : jf, kf, pk[PK_DL], tau);
      }
    }
  }
  double S[4] = {0., 0., 0., 0.};
  if (pol) {
    double E[4][4];
    Metric m;
    const double Xc[4] = {P.Xcam[0], P.Xcam[1], P.Xcam[2], P.Xcam[3]};
    camera_tetrad(G, Xc, 0, E, m);
    project(N, E, m, P.rotcam, S);
  }
  out7[0] = I;
  out7[1] = S[0];
  out7[2] = S[1];
  out7[3] = S[2];
  out7[4] = S[3];
  out7[5] = tau;
  out7[6] = tauF;
  return flag;
}

}  // namespace ipg
