/*
 * Per-ray bodies of the two passes, shared by the gfx950 kernels in
 * ipgpu.hip (one lane per ray).
 *
 * trace_ray:    the backward geodesic of one pixel, src/model_geodesics.c:38-212
 *               with push_photon (src/geodesics.c:9-64).  Writes one record of
 *               REC doubles per step: X1 X2 X3 K0 K1 K2 K3 dl (code units).
 * transfer_ray: integrate_emission (src/ipolarray.c:59-247) over those records,
 *               then project_N (:578-598).
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
  camera_tetrad(G, P.Xcam, P.old_centering, E, m);
  double X[4] = {P.Xcam[0], P.Xcam[1], P.Xcam[2], P.Xcam[3]};
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

/* out7 = Intensity, Is, Qs, Us, Vs, Tau, tauF (before the nu^3 scaling of
 * save_pixel); returns the oddflag bits. */
IPG_HD int transfer_ray(const ipgpu_geom &G, const ipgpu_model &Mo, const ipgpu_image &P, int ns,
                        const double *rec, const int *nt, const double *dump, double out7[7],
                        int &nrad)
{
  nrad = 0;
  const bool pol = !P.only_unpolarized;
  const double conv = Mo.L_unit * HPL / (ME * CL * CL);
  const double half_x2 = (G.cstopx[2] - G.cstartx[2]) / 2;

  Rec ti, tf, tn;
  load_rec(rec, ns, ti);
  Metric m;
  Fluid F;
  Coef c;
  double js, ks;
  metric(G, ti.X[1], ti.X[2], m);
  fluid(G, Mo, dump, ti.X, m, F);
  coefficients(F, ti.K, P.emission_type, false, js, ks, c);

  Herm N;
  for (int q = 0; q < 4; q++) N.d[q] = 0.;
  for (int q = 0; q < 6; q++) N.re[q] = N.im[q] = 0.;
  double I = 0., tau = 0., tauF = 0.;
  int flag = 0;
  bool Nnz = false, Mi_valid = false;
  double Mi[4][4];
  if (ns > 0) load_rec(rec, ns - 1, tn);
  for (int n = ns; n > 0; --n) {
    tf = tn;
    if (n > 1) load_rec(rec, n - 2, tn); /* prefetch the next record */
    double dl_cgs = ti.dl * Mo.L_unit * HPL / (ME * CL * CL);
    bool have_mf = false;
    if (pol && Nnz) {
      /* push_polar twice: src/ipolarray.c:85-89 */
      if (!Mi_valid) {
        metric(G, ti.X[1], ti.X[2], m);
        transport_matrix(m, ti.K, Mi);
      }
      Metric mh;
      double Xh[4], Kh[4], Mh[4][4];
      metric(G, tf.X[1], tf.X[2], m);
      have_mf = true;
      half_step(m, tf.X, tf.K, -ti.dl, Xh, Kh);
      metric(G, Xh[1], Xh[2], mh);
      transport_matrix(mh, Kh, Mh);
      Herm Nh;
      transport(N, N, Mi, (0.5 * dl_cgs) / conv, Nh);
      transport(N, Nh, Mh, dl_cgs / conv, N);
    }
    double r, th;
    if (have_mf) {
      r = m.r;
      th = m.th;
    } else {
      bl_coord(G, tf.X[1], tf.X[2], r, th);
    }
    if (radiating_region(G, Mo, r, th)) {
      nrad++;
      if (!have_mf) {
        metric(G, tf.X[1], tf.X[2], m);
        have_mf = true;
      }
      int zero_emission = 0;
      if (P.target_nturns >= 0 && nt && nt[n] != P.target_nturns) zero_emission = 1;
      if (P.counterjet == 1) {
        if (tf.X[2] < half_x2) zero_emission = 1;
      } else if (P.counterjet == 2) {
        if (tf.X[2] > half_x2) zero_emission = 1;
      }
      fluid(G, Mo, dump, tf.X, m, F);
      double jf, kf;
      coefficients(F, tf.K, P.emission_type, pol, jf, kf, c);
      double ji = js, ki = ks;
      js = jf;
      ks = kf;
      if (zero_emission) I = approximate_solve(I, 0, ki, 0, kf, dl_cgs, tau);
      else I = approximate_solve(I, ji, ki, jf, kf, dl_cgs, tau);
      if (pol) {
        flag |= evolve(F, m, tf.K, c, dl_cgs, zero_emission, P.stokes_floors, N, tauF);
        Nnz = herm_nonzero(N);
      }
    }
    if (pol && Nnz) {
      transport_matrix(m, tf.K, Mi); /* m holds the metric at tf.X here */
      Mi_valid = true;
    } else {
      Mi_valid = false;
    }
    ti = tf;
  }
  double S[4] = {0., 0., 0., 0.};
  if (pol) {
    double E[4][4];
    camera_tetrad(G, P.Xcam, 0, E, m);
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
