/*
 * Oracle (test infrastructure): geodesic integration, polarized transfer and
 * the image loop.  Restates src/geodesics.c, src/model_geodesics.c,
 * src/ipolarray.c and the fast-light part of src/main.c.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

extern or_geom_t OR_G;
extern or_model_t OR_M;

#define CUT_SMALL_OPTICAL_DEPTH 1e-5
#define CUT_PREVENT_NAN 1e-80

/* -------------------------------------------------------------- geodesics */

/* src/geodesics.c:9-64 */
void or_push_photon(double X[4], double Kcon[4], double dl, double Xhalf[4], double Kconhalf[4])
{
  double lconn[4][4][4], dKcon[4], Xh[4], Kconh[4];
  int i, j, k;
  or_connection(X, lconn);
  for (k = 0; k < 4; k++) dKcon[k] = 0.;
  for (i = 0; i < 4; i++)
    for (j = 0; j < 4; j++)
      for (k = 0; k < 4; k++) dKcon[k] -= 0.5 * dl * lconn[k][i][j] * Kcon[i] * Kcon[j];
  for (k = 0; k < 4; k++) Kconh[k] = Kcon[k] + dKcon[k];
  for (i = 0; i < 4; i++) Xh[i] = X[i] + 0.5 * dl * Kcon[i];
  for (i = 0; i < 4; i++) Xhalf[i] = Xh[i];
  for (i = 0; i < 4; i++) Kconhalf[i] = Kconh[i];
  or_connection(Xh, lconn);
  for (k = 0; k < 4; k++) dKcon[k] = 0.;
  for (i = 0; i < 4; i++)
    for (j = 0; j < 4; j++)
      for (k = 0; k < 4; k++) dKcon[k] -= dl * lconn[k][i][j] * Kconh[i] * Kconh[j];
  for (k = 0; k < 4; k++) Kcon[k] += dKcon[k];
  for (k = 0; k < 4; k++) X[k] += dl * Kconh[k];
}

/* src/model_geodesics.c:218-254 (no thin disk) */
int or_stop_backward(const double X[4], const double Kcon[4])
{
  double r, th;
  or_bl_coord(X, &r, &th);
  if ((r > OR_G.rmax_geo && Kcon[1] < 0.) || r < (OR_G.Rh + 0.0001)) return 1;
  return 0;
}

/* src/model_geodesics.c:267-293 (STEP_STRICT_MIN = 0) */
double or_stepsize(const double X[4], const double Kcon[4], double eps)
{
  const double SS = OR_SMALL * OR_SMALL;
  double deh = fmin(fabs(X[1] - OR_G.cstartx[1]), 0.1);
  double dlx1 = eps * (10 * deh) / (fabs(Kcon[1]) + SS);
  double cut = 0.02;
  double lx2 = OR_G.cstopx[2] - OR_G.cstartx[2];
  double dpole = fmin(fabs(X[2] / lx2), fabs((OR_G.cstopx[2] - X[2]) / lx2));
  double d2fac = (dpole < cut) ? dpole / 3 : fmin(cut / 3 + (dpole - cut) * 10., 1);
  double dlx2 = eps * d2fac / (fabs(Kcon[2]) + SS);
  double dlx3 = eps / (fabs(Kcon[3]) + SS);
  double idlx1 = 1. / (fabs(dlx1) + SS);
  double idlx2 = 1. / (fabs(dlx2) + SS);
  double idlx3 = 1. / (fabs(dlx3) + SS);
  return 1. / (idlx1 + idlx2 + idlx3);
}

/* src/model_geodesics.c:38-171 */
int or_trace(const double Xi[4], const double Kconi[4], or_traj_t *traj, double eps,
             int step_max, const double Xcam[4], int subring_dtheta)
{
  double X[4], Kcon[4], Xhalf[4], Kconhalf[4];
  for (int mu = 0; mu < 4; mu++) {
    X[mu] = Xi[mu];
    Xhalf[mu] = Xi[mu];
    Kcon[mu] = Kconi[mu];
    Kconhalf[mu] = Kconi[mu];
    traj[0].X[mu] = Xi[mu];
    traj[0].Kcon[mu] = Kconi[mu];
  }
  int nstep = 0, nturns = 0, passed_midplane_if_zero;
  if (Xcam[2] < 0.5) passed_midplane_if_zero = 1;
  else if (Xcam[2] > 0.5) passed_midplane_if_zero = -1;
  else passed_midplane_if_zero = 0;
  while ((!or_stop_backward(X, Kcon)) && (nstep < step_max - 1)) {
    nstep++;
    double dl = or_stepsize(X, Kcon, eps);
    traj[nstep].dl = dl * OR_M.L_unit * OR_HPL / (OR_ME * OR_CL * OR_CL);
    or_push_photon(X, Kcon, -dl, Xhalf, Kconhalf);
    for (int mu = 0; mu < 4; mu++) {
      traj[nstep].X[mu] = X[mu];
      traj[nstep].Kcon[mu] = Kcon[mu];
      traj[nstep].Xhalf[mu] = Xhalf[mu];
      traj[nstep].Kconhalf[mu] = Kconhalf[mu];
    }
    if (passed_midplane_if_zero != 0) {
      if (passed_midplane_if_zero > 0) {
        if (traj[nstep].X[2] > 0.5) passed_midplane_if_zero = 0;
      } else {
        if (traj[nstep].X[2] < 0.5) passed_midplane_if_zero = 0;
      }
    } else if (nstep > 3) {
      if (subring_dtheta == 1) {
        double dX2a = traj[nstep - 1].X[2] - traj[nstep - 2].X[2];
        double dX2b = traj[nstep].X[2] - traj[nstep - 1].X[2];
        if (dX2a * dX2b < 0) nturns++;
      } else {
        double ra, rb, rc, ha, hb, hc;
        or_bl_coord(traj[nstep - 2].X, &ra, &ha);
        or_bl_coord(traj[nstep - 1].X, &rb, &hb);
        or_bl_coord(traj[nstep - 0].X, &rc, &hc);
        double dz_1 = rb * cos(hb) - ra * cos(ha);
        double dz_2 = rc * cos(hc) - rb * cos(hb);
        if (dz_1 * dz_2 < 0) nturns++;
      }
    }
    traj[nstep].nturns = nturns;
  }
  return nstep;
}

/* src/geometry.c:190-201 */
static void null_normalize(double K[4], double fnorm)
{
  double inorm = sqrt(K[1] * K[1] + K[2] * K[2] + K[3] * K[3]);
  K[0] = fnorm;
  K[1] *= fnorm / inorm;
  K[2] *= fnorm / inorm;
  K[3] *= fnorm / inorm;
}

/* src/model_geodesics.c:178-212 and src/tetrads.c:22-33 */
void or_init_XK(long i, long j, const or_params_t *p, const or_camera_t *cam, double X[4],
                double Kcon[4])
{
  double Econ[4][4], Ecov[4][4], Kt[4];
  if (p->old_centering) or_camera_tetrad_old(cam->Xcam, Econ, Ecov);
  else or_camera_tetrad(cam->Xcam, Econ, Ecov);
  double dxoff = (i + 0.5 + p->xoff - 0.01) / p->nx - 0.5;
  double dyoff = (j + 0.5 + p->yoff) / p->ny - 0.5;
  Kt[0] = 0.;
  Kt[1] = (dxoff * cos(cam->rotcam_rad) - dyoff * sin(cam->rotcam_rad)) * cam->fovx;
  Kt[2] = (dxoff * sin(cam->rotcam_rad) + dyoff * cos(cam->rotcam_rad)) * cam->fovy;
  Kt[3] = 1.;
  null_normalize(Kt, 1.);
  for (int l = 0; l < 4; l++)
    Kcon[l] = Econ[0][l] * Kt[0] + Econ[1][l] * Kt[1] + Econ[2][l] * Kt[2] + Econ[3][l] * Kt[3];
  for (int mu = 0; mu < 4; mu++) X[mu] = cam->Xcam[mu];
}

/* -------------------------------------------------------------- transfer */

/* src/ipolarray.c:607-628 */
double or_approximate_solve(double Ii, double ji, double ki, double jf, double kf, double dl,
                            double *tau)
{
  double javg = (ji + jf) / 2.;
  double kavg = (ki + kf) / 2.;
  double dtau = dl * kavg;
  *tau += dtau;
  if (dtau < 1.e-3)
    return Ii + (javg - Ii * kavg) * dl * (1. - (dtau / 2.) * (1. - dtau / 3.));
  double efac = exp(-dtau);
  return Ii * efac + (javg / kavg) * (1. - efac);
}

/* src/ipolarray.c:253-289 */
void or_push_polar(const double Xm[4], const double Km[4], double complex Ni[4][4],
                   double complex Nm[4][4], double complex Nf[4][4], double dlam)
{
  double dl = dlam / (OR_M.L_unit * OR_HPL / (OR_ME * OR_CL * OR_CL));
  double lconn[4][4][4];
  or_connection(Xm, lconn);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) Nf[i][j] = Ni[i][j];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++)
        for (int l = 0; l < 4; l++)
          Nf[i][j] += -(lconn[i][k][l] * Nm[k][j] * Km[l] + lconn[j][k][l] * Nm[i][k] * Km[l]) * dl;
}

/* src/ipolarray.c:673-694 */
static void stokes_to_tensor(double fI, double fQ, double fU, double fV, double complex f[4][4])
{
  f[1][1] = (fI + fQ + 0. * I);
  f[1][2] = (fU - I * fV);
  f[2][1] = (fU + I * fV);
  f[2][2] = (fI - fQ + 0. * I);
}
static void tensor_to_stokes(double complex f[4][4], double *fI, double *fQ, double *fU, double *fV)
{
  *fI = creal(f[1][1] + f[2][2]) / 2;
  *fQ = creal(f[1][1] - f[2][2]) / 2;
  *fU = creal(f[1][2] + f[2][1]) / 2;
  *fV = cimag(f[2][1] - f[1][2]) / 2;
}

/* src/ipolarray.c:710-748 */
static void coord_to_tetrad_rank2(double complex Tc[4][4], double Ecov[4][4], double complex Tt[4][4])
{
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) Tt[i][j] = 0. + I * 0.;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++)
        for (int l = 0; l < 4; l++) Tt[i][j] += Tc[k][l] * Ecov[i][k] * Ecov[j][l];
}
static void tetrad_to_coord_rank2(double complex Tt[4][4], double Econ[4][4], double complex Tc[4][4])
{
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) Tc[i][j] = 0. + I * 0.;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++)
        for (int l = 0; l < 4; l++) Tc[i][j] += Tt[k][l] * Econ[k][i] * Econ[l][j];
}

/* The Stokes update of evolve_N in the plasma frame, src/ipolarray.c:390-520:
 * Faraday half step, absorption/emission, Faraday half step.  S = I Q U V. */
void or_stokes_step(const double c11[11], double dlam, double S[4])
{
  double jI = c11[0], jQ = c11[1], jU = c11[2], jV = c11[3];
  double aI = c11[4], aQ = c11[5], aU = c11[6], aV = c11[7];
  double rQ = c11[8], rU = c11[9], rV = c11[10];
  double SI0 = S[0], SQ0 = S[1], SU0 = S[2], SV0 = S[3];
  double SI1, SQ1, SU1, SV1, SI2, SQ2, SU2, SV2, SI, SQ, SU, SV;
  double x = dlam * 0.5;
  double rho2 = rQ * rQ + rU * rU + rV * rV;
  if (rho2 > CUT_PREVENT_NAN) {
    double rho = sqrt(rho2);
    double rdS = rQ * SQ0 + rU * SU0 + rV * SV0;
    double c = cos(rho * x), s = sin(rho * x), sh = sin(0.5 * rho * x);
    SI1 = SI0;
    SQ1 = SQ0 * c + 2 * rQ * rdS / rho2 * sh * sh + (rU * SV0 - rV * SU0) / rho * s;
    SU1 = SU0 * c + 2 * rU * rdS / rho2 * sh * sh + (rV * SQ0 - rQ * SV0) / rho * s;
    SV1 = SV0 * c + 2 * rV * rdS / rho2 * sh * sh + (rQ * SU0 - rU * SQ0) / rho * s;
  } else {
    SI1 = SI0;
    SQ1 = SQ0 + (-rV * SU0 + rU * SV0) * x;
    SU1 = SU0 + (rV * SQ0 - rQ * SV0) * x;
    SV1 = SV0 + (-rU * SQ0 + rQ * SU0) * x;
  }

  x = dlam;
  double aP2 = aQ * aQ + aU * aU + aV * aV;
  if (aP2 > CUT_PREVENT_NAN) {
    double aP = sqrt(aP2);
    double tauP = aP * x;
    double tauI = aI * x;
    double ads0 = aQ * SQ1 + aU * SU1 + aV * SV1;
    double adj = aQ * jQ + aU * jU + aV * jV;
    double efacm = exp(-tauI + tauP);
    double efacp = exp(-tauI - tauP);
    double efac = exp(-tauI);
    SI2 = efacm * (SI1 / 2. - ads0 / (2. * aP)) + efacp * (SI1 / 2. + ads0 / (2. * aP));
    SQ2 = efacm * (-SI1 * aQ * aP + ads0 * aQ) / (2. * aP2)
        + efacp * (SI1 * aQ * aP + ads0 * aQ) / (2. * aP2) + efac * (SQ1 - ads0 * aQ / (aP2));
    SU2 = efacm * (-SI1 * aU * aP + ads0 * aU) / (2. * aP2)
        + efacp * (SI1 * aU * aP + ads0 * aU) / (2. * aP2) + efac * (SU1 - ads0 * aU / (aP2));
    SV2 = efacm * (-SI1 * aV * aP + ads0 * aV) / (2. * aP2)
        + efacp * (SI1 * aV * aP + ads0 * aV) / (2. * aP2) + efac * (SV1 - ads0 * aV / (aP2));
    double afacm = 1. - efacm, afac = 1. - efac, afacp = 1. - efacp;
    if (tauI - tauP <= CUT_SMALL_OPTICAL_DEPTH) {
      double e = tauI - tauP;
      afacm = e * (1. - (e / 2.) * (1. - e / 3.));
      if (tauI <= CUT_SMALL_OPTICAL_DEPTH) {
        e = tauI;
        afac = e * (1. - (e / 2.) * (1. - e / 3.));
        if (tauI + tauP <= CUT_SMALL_OPTICAL_DEPTH) {
          e = tauI + tauP;
          afacp = e * (1. - (e / 2.) * (1. - e / 3.));
        }
      }
    }
    SQ2 += afac * (jQ / aI - aQ * adj / (aI * aP2));
    SU2 += afac * (jU / aI - aU * adj / (aI * aP2));
    SV2 += afac * (jV / aI - aV * adj / (aI * aP2));
    SI2 += afacm * (aP * jI - adj) / (2. * aP * (aI - aP));
    SQ2 += afacm * aQ * (-aP * jI + adj) / (2. * aP2 * (aI - aP));
    SU2 += afacm * aU * (-aP * jI + adj) / (2. * aP2 * (aI - aP));
    SV2 += afacm * aV * (-aP * jI + adj) / (2. * aP2 * (aI - aP));
    SI2 += afacp * (aP * jI + adj) / (2. * aP * (aI + aP));
    SQ2 += afacp * aQ * (aP * jI + adj) / (2. * aP2 * (aI + aP));
    SU2 += afacp * aU * (aP * jI + adj) / (2. * aP2 * (aI + aP));
    SV2 += afacp * aV * (aP * jI + adj) / (2. * aP2 * (aI + aP));
  } else {
    double tau_fake = 0;
    SI2 = or_approximate_solve(SI1, jI, aI, jI, aI, x, &tau_fake);
    SQ2 = or_approximate_solve(SQ1, jQ, aI, jQ, aI, x, &tau_fake);
    SU2 = or_approximate_solve(SU1, jU, aI, jU, aI, x, &tau_fake);
    SV2 = or_approximate_solve(SV1, jV, aI, jV, aI, x, &tau_fake);
  }

  x = dlam * 0.5;
  rho2 = rQ * rQ + rU * rU + rV * rV;
  if (rho2 > CUT_PREVENT_NAN) {
    double rdS = rQ * SQ2 + rU * SU2 + rV * SV2;
    double rho = sqrt(rho2);
    double c = cos(rho * x), s = sin(rho * x), sh = sin(0.5 * rho * x);
    SI = SI2;
    SQ = SQ2 * c + 2 * rQ * rdS / rho2 * sh * sh + (rU * SV2 - rV * SU2) / rho * s;
    SU = SU2 * c + 2 * rU * rdS / rho2 * sh * sh + (rV * SQ2 - rQ * SV2) / rho * s;
    SV = SV2 * c + 2 * rV * rdS / rho2 * sh * sh + (rQ * SU2 - rU * SQ2) / rho * s;
  } else {
    SI = SI2;
    SQ = SQ2 + (-rV * SU2 + rU * SV2) * x;
    SU = SU2 + (rV * SQ2 - rQ * SV2) * x;
    SV = SV2 + (-rU * SQ2 + rQ * SU2) * x;
  }
  S[0] = SI;
  S[1] = SQ;
  S[2] = SU;
  S[3] = SV;
}

/* src/ipolarray.c:332-575 */
int or_evolve_N(const double Xf[4], const double Kconf[4], double dlam, double complex N_coord[4][4],
                double *tauF, int zero_emission, const or_params_t *p)
{
  double gcov[4][4], Ucon[4], Bcon[4], Ucov[4], Bcov[4], Ecov[4][4], Econ[4][4];
  double complex N_tetrad[4][4];
  double c11[11];
  double SI0, SQ0, SU0, SV0, SI, SQ, SU, SV;
  int oddflag = 0;
  or_jar(p->emission_type, 1, Xf, Kconf, c11);
  double jI = c11[0], jQ = c11[1], jU = c11[2], jV = c11[3];
  double aI = c11[4], aQ = c11[5], aU = c11[6], aV = c11[7];
  double rQ = c11[8], rU = c11[9], rV = c11[10];
  if (zero_emission) {
    jI = 0.;
    jQ = 0.;
    jU = 0.;
    jV = 0.;
  }
  or_fourv(Xf, Kconf, Ucon, Ucov, Bcon, Bcov);
  double bsq = 0.;
  for (int mu = 0; mu < 4; mu++) bsq += Bcon[mu] * Bcov[mu];
  if (bsq <= 0.) {
    Bcon[0] = 0.;
    Bcon[1] = 1.;
    Bcon[2] = 1.;
    Bcon[3] = 1.;
  }
  or_gcov(Xf, gcov);
  oddflag |= or_plasma_tetrad(Ucon, Kconf, Bcon, gcov, Econ, Ecov);
  coord_to_tetrad_rank2(N_coord, Ecov, N_tetrad);
  tensor_to_stokes(N_tetrad, &SI0, &SQ0, &SU0, &SV0);

  double S4[4] = {SI0, SQ0, SU0, SV0};
  double c11s[11] = {jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV};
  or_stokes_step(c11s, dlam, S4);
  SI = S4[0];
  SQ = S4[1];
  SU = S4[2];
  SV = S4[3];
  *tauF += dlam * fabs(rV);
  if (p->stokes_floors) {
    if (SI < 0) {
      SI = 0;
      SQ = 0;
      SU = 0;
      SV = 0;
    } else {
      double pol_frac = sqrt(SQ * SQ + SU * SU + SV * SV) / SI;
      if (pol_frac > 1.) {
        SQ /= pol_frac;
        SU /= pol_frac;
        SV /= pol_frac;
      }
    }
  }
  stokes_to_tensor(SI, SQ, SU, SV, N_tetrad);
  tetrad_to_coord_rank2(N_tetrad, Econ, N_coord);
  if (*tauF > 1.e100 || *tauF < -1.e100 || isnan(*tauF)) oddflag |= 2;
  if (isnan(creal(N_tetrad[0][0])) || isnan(creal(N_coord[0][0]))) oddflag |= 4;
  if (isnan(Ucov[0])) oddflag |= 8;
  return oddflag;
}

/* src/ipolarray.c:578-598 */
void or_project_N(const double X[4], double complex N_coord[4][4], double *Stokes_I,
                  double *Stokes_Q, double *Stokes_U, double *Stokes_V, double rotcam)
{
  double complex N_tetrad[4][4];
  double Econ[4][4], Ecov[4][4], Q, U;
  or_camera_tetrad(X, Econ, Ecov);
  coord_to_tetrad_rank2(N_coord, Ecov, N_tetrad);
  tensor_to_stokes(N_tetrad, Stokes_I, &Q, &U, Stokes_V);
  rotcam *= -2.;
  *Stokes_Q = Q * cos(rotcam) - U * sin(rotcam);
  *Stokes_U = Q * sin(rotcam) + U * cos(rotcam);
}

/* src/model_radiation.c:119-123 */
static void get_jkinv(const double X[4], const double K[4], double *j, double *k, int dist)
{
  double c[11];
  or_jar(dist, 0, X, K, c);
  *j = c[0];
  *k = c[4];
}

/* src/ipolarray.c:59-247 (non-debug build: flags are OR-ed) */
int or_integrate_emission(or_traj_t *traj, int nsteps, double *Intensity, double *Tau,
                          double *tauF, double complex N_coord[4][4], const or_params_t *p)
{
  *tauF = 0.;
  *Intensity = 0.;
  *Tau = 0.;
  double js = 0., ks;
  int oddflag = 0;
  get_jkinv(traj[nsteps].X, traj[nsteps].Kcon, &js, &ks, p->emission_type);
  for (int nstep = nsteps; nstep > 0; --nstep) {
    int sflag = 0;
    or_traj_t ti = traj[nstep];
    or_traj_t tf = traj[nstep - 1];
    if (!p->only_unpolarized) {
      double complex Nh[4][4];
      or_push_polar(ti.X, ti.Kcon, N_coord, N_coord, Nh, 0.5 * ti.dl);
      or_push_polar(ti.Xhalf, ti.Kconhalf, N_coord, Nh, N_coord, ti.dl);
    }
    if (or_radiating_region(tf.X)) {
      int zero_emission = 0;
      if (p->target_nturns >= 0 && ti.nturns != p->target_nturns) zero_emission = 1;
      if (p->counterjet == 1) {
        if (tf.X[2] < (OR_G.cstopx[2] - OR_G.cstartx[2]) / 2) zero_emission = 1;
      } else if (p->counterjet == 2) {
        if (tf.X[2] > (OR_G.cstopx[2] - OR_G.cstartx[2]) / 2) zero_emission = 1;
      }
      double ji = js, ki = ks, jf, kf;
      get_jkinv(tf.X, tf.Kcon, &jf, &kf, p->emission_type);
      js = jf;
      ks = kf;
      if (zero_emission) *Intensity = or_approximate_solve(*Intensity, 0, ki, 0, kf, ti.dl, Tau);
      else *Intensity = or_approximate_solve(*Intensity, ji, ki, jf, kf, ti.dl, Tau);
      if (!p->only_unpolarized)
        sflag |= or_evolve_N(tf.X, tf.Kcon, ti.dl, N_coord, tauF, zero_emission, p);
    }
    oddflag |= sflag;
  }
  return oddflag;
}

/* -------------------------------------------------------------- image */

/* src/main.c:125-174 */
int or_setup_camera(const or_params_t *p, or_camera_t *cam)
{
  memset(cam, 0, sizeof(*cam));
  cam->freq = p->freqcgs * OR_HPL / (OR_ME * OR_CL * OR_CL);
  if (cam->freq == 0) return -1;
  cam->rotcam_rad = p->rotcam * M_PI / 180.;
  or_native_coord(p->rcam, p->thetacam, p->phicam, cam->Xcam);
  cam->dsource_cm = p->dsource * OR_PC;
  double Dsource = cam->dsource_cm;
  double L_unit = OR_M.L_unit;
  double fov_to_d = Dsource / L_unit / OR_MUAS_PER_RAD;
  double fx = p->fovx_dsource, fy = p->fovy_dsource;
  if (fx != 0.0) {
  } else if (p->dx != 0.0) {
    fx = p->dx / fov_to_d;
    fy = p->dy / fov_to_d;
  } else {
    fx = 160.;
    fy = 160.;
  }
  cam->DX = fx * fov_to_d;
  cam->DY = fy * fov_to_d;
  cam->fovx = cam->DX / p->rcam;
  cam->fovy = cam->DY / p->rcam;
  cam->scale = (cam->DX * L_unit / p->nx) * (cam->DY * L_unit / p->ny) / (Dsource * Dsource) / OR_JY;
  return 0;
}

/* One pixel: src/main.c:822-883 (get_pixel) without trace output.
 * out7 = Intensity, Is, Qs, Us, Vs, Tau, tauF (before the nu^3 scaling). */
static int pixel(long i, long j, const or_params_t *p, const or_camera_t *cam, or_traj_t *traj,
                 double out7[7], int *nstep_out, int *flag_out)
{
  double X[4] = {0.}, Kcon[4] = {0.};
  double complex N[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) N[a][b] = 0.;
  or_init_XK(i, j, p, cam, X, Kcon);
  memset(traj, 0, sizeof(or_traj_t)); /* traj[0]; later entries are written before use */
  for (int mu = 0; mu < 4; mu++) Kcon[mu] *= cam->freq;
  int nstep = or_trace(X, Kcon, traj, p->eps, p->maxnstep, cam->Xcam, p->subring_dtheta);
  *nstep_out = nstep;
  if (nstep >= p->maxnstep - 1) return -10;
  double Intensity, Tau, tauF, Is = 0, Qs = 0, Us = 0, Vs = 0;
  int oddflag = or_integrate_emission(traj, nstep, &Intensity, &Tau, &tauF, N, p);
  if (!p->only_unpolarized) or_project_N(X, N, &Is, &Qs, &Us, &Vs, cam->rotcam_rad);
  *flag_out = oddflag;
  out7[0] = Intensity;
  out7[1] = Is;
  out7[2] = Qs;
  out7[3] = Us;
  out7[4] = Vs;
  out7[5] = Tau;
  out7[6] = tauF;
  return 0;
}

/* ------------------------------------------------------ adaptive refinement */

/* save_pixel, src/main.c:885-910, into an image of row length ny */
static void save_px(double *image, double *imageS, double *taus, size_t i, size_t j, size_t ny,
                    const or_params_t *p, const double o[7])
{
  const double f3 = pow(p->freqcgs, 3);
  image[i * ny + j] = o[0] * f3;
  taus[i * ny + j] = o[5];
  if (!p->only_unpolarized) {
    double *S = imageS + (i * ny + j) * 5;
    S[0] = o[1] * f3;
    S[1] = (p->qu_conv == 0 ? -o[2] : o[2]) * f3;
    S[2] = (p->qu_conv == 0 ? -o[3] : o[3]) * f3;
    S[3] = o[4] * f3;
    S[4] = o[6];
  }
}

/* Adaptive refinement, src/main.c:98-124 (levels and the internal 2^N + 1
 * grid) and :505-792 (the coarse pass, the intensity estimate and the
 * levels), with save_pixelTransfer, lininterp2 and lininterp4 (:912-1000).
 * The pixels a level traces are decided first and traced together: each
 * decision reads only the coarser grid, so this equals the reference's
 * interleaved loop.  Outputs are the nx x ny part of the internal grid;
 * Iavg, traced[level] and interpolated[level] (the running count the
 * reference prints) as ipgpu_adaptive_stats reports them.  Returns 0, -10
 * (maxnstep), or -2 when the sizes do not form the 2^N + 1 scheme. */
int or_render_adaptive(const or_params_t *p, or_camera_t *cam, int nx_min, int ny_min, double refine_abs,
                       double refine_rel, double refine_cut, int nearest, double *image_out,
                       double *imageS_out, double *taus_out, int *interp_out, double *Iavg_out,
                       long long *traced, long long *interpolated, int nthreads)
{
  int rc = or_setup_camera(p, cam);
  if (rc) return rc;
  const int refine_level = (int)(log2((p->nx - p->nx % 2) / (nx_min - nx_min % 2)) + 1);
  size_t nx, ny, nxmin, nymin;
  if (refine_level > 1 && p->nx % 2 == 0) {
    nx = p->nx + 1;
    nxmin = nx_min + 1;
  } else {
    nx = p->nx;
    nxmin = nx_min;
  }
  if (refine_level > 1 && p->ny % 2 == 0) {
    ny = p->ny + 1;
    nymin = ny_min + 1;
  } else {
    ny = p->ny;
    nymin = ny_min;
  }
  const size_t sp = (nx - 1) / (nxmin - 1), spy = (ny - 1) / (nymin - 1);
  if (refine_level < 2 || sp != ((size_t)1 << (refine_level - 1)) || spy != sp || (nx - 1) % (nxmin - 1)
      || (ny - 1) % (nymin - 1))
    return -2;
  double *image = calloc(nx * ny, sizeof(double)), *taus = calloc(nx * ny, sizeof(double));
  double *imageS = calloc(nx * ny * 5, sizeof(double));
  int *interp = calloc(nx * ny, sizeof(int));
  int *pix = malloc(sizeof(int) * 2 * nx * ny);
  double *out7 = malloc(sizeof(double) * 7 * nx * ny);
  int *ns = malloc(sizeof(int) * nx * ny), *fl = malloc(sizeof(int) * nx * ny);
  double *prelim = calloc(nxmin * nymin, sizeof(double));
  /* coarse pass */
  int n = 0;
  for (size_t i = 0; i < nx; i += sp)
    for (size_t j = 0; j < ny; j += sp) {
      pix[2 * n] = (int)i;
      pix[2 * n + 1] = (int)j;
      n++;
    }
  rc = or_render_pixels(p, cam, pix, n, out7, ns, fl, nthreads);
  if (rc == 0) {
    for (int q = 0; q < n; q++) {
      const size_t i = pix[2 * q], j = pix[2 * q + 1];
      save_px(image, imageS, taus, i, j, ny, p, out7 + 7 * (size_t)q);
      const double Intensity = out7[7 * (size_t)q];
      const size_t loc = i / sp * nymin + j / sp;
      if (i % (nx - 1) != 0 && j % (ny - 1) != 0) {
        prelim[loc] = Intensity * sp * sp;
      } else if ((i == 0 && j % (ny - 1) != 0) || (i % (nx - 1) != 0 && j == 0)) {
        prelim[loc] = Intensity * (sp / 2 + 1) * sp;
      } else if (i % (nx - 1) != 0 || j % (ny - 1) != 0) {
        prelim[loc] = Intensity * (sp / 2) * sp;
      } else {
        if (i == 0 && j == 0) prelim[loc] = Intensity * (sp / 2 + 1) * (sp / 2 + 1);
        else if (i == 0 || j == 0) prelim[loc] = Intensity * (sp / 2 + 1) * sp / 2;
        else prelim[loc] = Intensity * (sp * sp) / 4;
      }
    }
    traced[0] = n;
    interpolated[0] = 0;
  }
  double interp_tot = 0;
  for (size_t q = 0; q < nxmin * nymin; q++) interp_tot += prelim[q];
  const double Iavg = interp_tot * pow(p->freqcgs, 3) / (nx * ny);
  *Iavg_out = Iavg;
  long long total_interp = 0;
  for (int lev = 1; lev < refine_level && rc == 0; lev++) {
    const size_t nsx = sp / (size_t)pow(2, lev), prev = nsx * 2;
    n = 0;
    for (size_t i = 0; i < nx; i += nsx)
      for (size_t j = 0; j < ny; j += nsx) {
        double I1, I2, I3, err_abs, err_rel;
        size_t s0, s1, s2 = 0, s3 = 0;
        int four = 0;
        if (i % prev == 0 && j % prev == 0) continue;
        if (i % prev == 0) {
          s0 = i * ny + j - nsx;
          s1 = i * ny + j + nsx;
          I1 = image[s0];
          I2 = image[s1];
          err_abs = (I2 - I1) / 2 / Iavg;
          err_rel = (I2 - I1) / 2 / I1;
        } else if (j % prev == 0) {
          s0 = (i - nsx) * ny + j;
          s1 = (i + nsx) * ny + j;
          I1 = image[s0];
          I2 = image[s1];
          err_abs = (I2 - I1) / 2 / Iavg;
          err_rel = (I2 - I1) / 2 / I1;
        } else {
          s0 = (i - nsx) * ny + j - nsx;
          s1 = (i + nsx) * ny + j - nsx;
          s2 = (i - nsx) * ny + j + nsx;
          s3 = (i + nsx) * ny + (j + nsx);
          four = 1;
          I1 = image[s0];
          I2 = image[s1];
          I3 = image[s2];
          err_abs = ((I2 + I3) / 2 - I1) / Iavg;
          err_rel = (I2 + I3) / (2 * I1) - 1.;
        }
        const size_t t = i * ny + j;
        if ((fabs(err_abs) > refine_abs && fabs(err_rel) > refine_rel) && fabs(I1) > refine_cut) {
          pix[2 * n] = (int)i;
          pix[2 * n + 1] = (int)j;
          n++;
          interp[t] = 0;
          continue;
        }
        interp[t] = 1;
        total_interp++;
        if (nearest) {
          image[t] = image[s0];
          if (!p->only_unpolarized) {
            taus[t] = taus[s0];
            for (int k = 0; k < 5; k++) imageS[t * 5 + k] = imageS[s0 * 5 + k];
          }
        } else if (!four) {
          image[t] = .5 * (image[s0] + image[s1]);
          taus[t] = .5 * (taus[s0] + taus[s1]);
          if (!p->only_unpolarized)
            for (int k = 0; k < 5; k++) imageS[t * 5 + k] = .5 * (imageS[s0 * 5 + k] + imageS[s1 * 5 + k]);
        } else {
          image[t] = .25 * (image[s0] + image[s1] + image[s2] + image[s3]);
          taus[t] = .25 * (taus[s0] + taus[s1] + taus[s2] + taus[s3]);
          if (!p->only_unpolarized)
            for (int k = 0; k < 5; k++)
              imageS[t * 5 + k] = .25 * (imageS[s0 * 5 + k] + imageS[s1 * 5 + k] + imageS[s2 * 5 + k]
                                         + imageS[s3 * 5 + k]);
        }
      }
    if (n) rc = or_render_pixels(p, cam, pix, n, out7, ns, fl, nthreads);
    if (rc == 0)
      for (int q = 0; q < n; q++) save_px(image, imageS, taus, pix[2 * q], pix[2 * q + 1], ny, p, out7 + 7 * (size_t)q);
    traced[lev] = n;
    interpolated[lev] = total_interp;
  }
  for (size_t i = 0; i < (size_t)p->nx; i++)
    for (size_t j = 0; j < (size_t)p->ny; j++) {
      const size_t a = i * p->ny + j, b = i * ny + j;
      image_out[a] = image[b];
      taus_out[a] = taus[b];
      interp_out[a] = interp[b];
      if (!p->only_unpolarized)
        for (int k = 0; k < 5; k++) imageS_out[a * 5 + k] = imageS[b * 5 + k];
    }
  free(image);
  free(taus);
  free(imageS);
  free(interp);
  free(pix);
  free(out7);
  free(ns);
  free(fl);
  free(prelim);
  return rc;
}

/* src/ipolarray.c:297-320: N carried by dl (cgs) with the connection at
 * Xm, Km */
static void transport_vector(const double Xm[4], const double Km[4], const double Ni[4],
                             const double Nm[4], double Nf[4], double dl)
{
  double lconn[4][4][4];
  or_connection(Xm, lconn);
  for (int i = 0; i < 4; i++) Nf[i] = Ni[i];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++)
        Nf[i] += -(lconn[i][j][k] * Nm[j] * Km[k]) * dl / (OR_M.L_unit * OR_HPL / (OR_ME * OR_CL * OR_CL));
}

/* src/radiation.c:47-72 */
static double bk_angle(const double K[4], const double Ucov[4], const double Bcon[4],
                       const double Bcov[4])
{
  double B = sqrt(fabs(Bcon[0] * Bcov[0] + Bcon[1] * Bcov[1] + Bcon[2] * Bcov[2] + Bcon[3] * Bcov[3]));
  if (B == 0.) return M_PI / 2.;
  double k = fabs(K[0] * Ucov[0] + K[1] * Ucov[1] + K[2] * Ucov[2] + K[3] * Ucov[3]);
  double mu = (K[0] * Bcov[0] + K[1] * Bcov[1] + K[2] * Bcov[2] + K[3] * Bcov[3]) / (k * B);
  if (fabs(mu) > 1.) mu /= fabs(mu);
  return acos(mu);
}

/* Trace output of pixel (i, j): get_pixel (src/main.c:822-883) with
 * params.trace, then the per-step arrays of dump_var_along
 * (src/io.c:279-543) as records of OR_TR_NF doubles at the OR_TR_* offsets,
 * record s holding index s of every array (zero where the reference leaves
 * its calloc'ed arrays unwritten).  rub: the ghosted rho, u, beta arrays of
 * or_iharm_derive_ru / its beta, or NULL (those entries stay 0).  Returns 0,
 * -10 when maxnstep is exceeded, -1 when cap < nstep + 1. */
int or_trace_pixel(const or_params_t *p, const or_camera_t *cam, long i, long j, const double *rub,
                   int cap, double *out, int *nstep_out)
{
  double X[4] = {0.}, Kcon[4] = {0.};
  or_init_XK(i, j, p, cam, X, Kcon);
  for (int mu = 0; mu < 4; mu++) Kcon[mu] *= cam->freq;
  or_traj_t *traj = calloc((size_t)p->maxnstep, sizeof(or_traj_t));
  int nstep = or_trace(X, Kcon, traj, p->eps, p->maxnstep, cam->Xcam, p->subring_dtheta);
  *nstep_out = nstep;
  if (nstep >= p->maxnstep - 1) {
    free(traj);
    return -10;
  }
  if (cap < nstep + 1) {
    free(traj);
    return -1;
  }
  memset(out, 0, sizeof(double) * (size_t)cap * OR_TR_NF);
  double complex N_coord[4][4], N_tetrad[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) N_coord[a][b] = 0.;
  double Intensity = 0, Tau = 0, tauF = 0;
  for (int s = nstep - 1; s > 0; --s) {
    double *o = out + (size_t)s * OR_TR_NF;
    const or_traj_t *t = &traj[s];
    double Ucon[4], Ucov[4], Bcon[4], Bcov[4];
    or_primitives(t->X, rub, o + OR_TR_PRIMS, o + OR_TR_BETA);
    or_fourv(t->X, t->Kcon, Ucon, Ucov, Bcon, Bcov);
    o[OR_TR_B] = or_b(t->X);
    o[OR_TR_NE] = or_ne(t->X);
    o[OR_TR_THETAE] = or_thetae(t->X);
    /* get_fluid_nu, src/radiation.c:20-42 */
    o[OR_TR_NU] = -(t->Kcon[0] * Ucov[0] + t->Kcon[1] * Ucov[1] + t->Kcon[2] * Ucov[2]
                    + t->Kcon[3] * Ucov[3]) * OR_ME * OR_CL * OR_CL / OR_HPL;
    o[OR_TR_BK_ANGLE] = bk_angle(t->Kcon, Ucov, Bcon, Bcov);
    o[OR_TR_SIGMA] = or_sigma(t->X);
    double c[11];
    or_jar(p->emission_type, 1, t->X, t->Kcon, c);
    for (int q = 0; q < 4; q++) {
      o[OR_TR_J_INV + q] = c[q];
      o[OR_TR_ALPHA_INV + q] = c[4 + q];
    }
    for (int q = 0; q < 3; q++) o[OR_TR_RHO_INV + 1 + q] = c[8 + q];
    get_jkinv(t->X, t->Kcon, o + OR_TR_J_UNPOL, o + OR_TR_K_UNPOL, p->emission_type);
    o[OR_TR_DL] = t->dl;
    for (int mu = 0; mu < 4; mu++) {
      o[OR_TR_X + mu] = t->X[mu];
      o[OR_TR_KCON + mu] = t->Kcon[mu];
    }
    double g[4][4];
    or_gcov(t->X, g);
    or_bl_coord(t->X, o + OR_TR_R, o + OR_TR_TH);
    o[OR_TR_PHI] = t->X[3];
    double Econ[4][4], Ecov[4][4];
    or_plasma_tetrad(Ucon, t->Kcon, Bcon, g, Econ, Ecov);
    double e2Mid[4], e2Final[4];
    transport_vector(t->X, t->Kcon, Econ[2], Econ[2], e2Mid, t->dl);
    transport_vector(t->Xhalf, t->Kconhalf, Econ[2], e2Mid, e2Final, t->dl);
    double *om = out + (size_t)(s - 1) * OR_TR_NF;
    for (int mu = 0; mu < 4; mu++) {
      o[OR_TR_E2C + mu] = Econ[2][mu];
      o[OR_TR_E2C_COV + mu] = Ecov[2][mu];
      om[OR_TR_E2T + mu] = e2Final[mu];
      e2Final[mu] = o[OR_TR_E2T + mu];
    }
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        o[OR_TR_ECON + 4 * a + b] = Econ[a][b];
        o[OR_TR_ECOV + 4 * a + b] = Ecov[a][b];
      }
    or_lower(e2Final, g, o + OR_TR_E2T_COV);
    coord_to_tetrad_rank2(N_coord, Ecov, N_tetrad);
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        o[OR_TR_NTETRAD + 4 * a + b] = creal(N_tetrad[a][b]);
        o[OR_TR_NCOORD + 4 * a + b] = creal(N_coord[a][b]);
      }
    tensor_to_stokes(N_tetrad, o + OR_TR_STOKES, o + OR_TR_STOKES + 1, o + OR_TR_STOKES + 2,
                     o + OR_TR_STOKES + 3);
    or_integrate_emission(&traj[s], 1, &Intensity, &Tau, &tauF, N_coord, p);
    o[OR_TR_I_UNPOL] = Intensity;
  }
  free(traj);
  return 0;
}

/* Pixels listed as (i, j) pairs; returns 0 or -10 (maxnstep exceeded). */
int or_render_pixels(const or_params_t *p, const or_camera_t *cam, const int *pix, int npix,
                     double *out7, int *nsteps, int *flags, int nthreads)
{
  int status = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    or_traj_t *traj = malloc(sizeof(or_traj_t) * (size_t)p->maxnstep);
#pragma omp for schedule(dynamic, 1)
    for (int n = 0; n < npix; n++) {
      int rc = pixel(pix[2 * n], pix[2 * n + 1], p, cam, traj, out7 + 7 * (size_t)n, nsteps + n,
                     flags + n);
      if (rc) {
#pragma omp critical
        status = rc;
      }
    }
    free(traj);
  }
  return status;
}

/* src/main.c:536-615 (fast light, no refinement) and save_pixel :885-910 */
int or_render(const or_params_t *p, or_camera_t *cam, double *image, double *imageS, double *taus,
              int *nsteps, int *flags, int nthreads)
{
  int rc = or_setup_camera(p, cam);
  if (rc) return rc;
  size_t nx = p->nx, ny = p->ny, npix = nx * ny;
  int *pix = malloc(sizeof(int) * 2 * npix);
  double *out7 = malloc(sizeof(double) * 7 * npix);
  for (size_t i = 0; i < nx; i++)
    for (size_t j = 0; j < ny; j++) {
      pix[2 * (i * ny + j)] = (int)i;
      pix[2 * (i * ny + j) + 1] = (int)j;
    }
  rc = or_render_pixels(p, cam, pix, (int)npix, out7, nsteps, flags, nthreads);
  if (rc == 0) {
    double f3 = pow(p->freqcgs, 3);
    for (size_t n = 0; n < npix; n++) {
      double *o = out7 + 7 * n;
      image[n] = o[0] * f3;
      taus[n] = o[5];
      if (!p->only_unpolarized) {
        imageS[n * 5 + 0] = o[1] * f3;
        if (p->qu_conv == 0) {
          imageS[n * 5 + 1] = -o[2] * f3;
          imageS[n * 5 + 2] = -o[3] * f3;
        } else {
          imageS[n * 5 + 1] = o[2] * f3;
          imageS[n * 5 + 2] = o[3] * f3;
        }
        imageS[n * 5 + 3] = o[4] * f3;
        imageS[n * 5 + 4] = o[6];
        if (isnan(imageS[n * 5])) rc = -1;
      }
    }
  }
  free(pix);
  free(out7);
  return rc;
}
