/*
 * Oracle (test infrastructure): coordinates, metric, connection, tetrads.
 * Restates src/coordinates.c, src/geometry.c, src/tetrads.c and
 * src/model_tetrads.c of the reference for the MKS, FMKS, Minkowski and
 * eMinkowski systems.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

or_geom_t OR_G;

void or_set_geom(const or_geom_t *g) { OR_G = *g; }
void or_get_geom(or_geom_t *g) { *g = OR_G; }

/* src/coordinates.c:46-97 */
void or_bl_coord(const double X[4], double *r, double *th)
{
  if (OR_G.metric == OR_METRIC_MINKOWSKI) {
    *r = X[1];
    *th = X[2];
    return;
  } else if (OR_G.metric == OR_METRIC_EMINKOWSKI) {
    *r = exp(X[1]);
    *th = X[2];
    return;
  }
  *r = exp(X[1]);
  if (OR_G.metric == OR_METRIC_MKS) {
    *th = M_PI * X[2] + ((1. - OR_G.hslope) / 2.) * sin(2. * M_PI * X[2]);
  } else { /* FMKS */
    double thG = M_PI * X[2] + ((1. - OR_G.hslope) / 2.) * sin(2. * M_PI * X[2]);
    double y = 2 * X[2] - 1.;
    double thJ = OR_G.poly_norm * y
        * (1. + pow(y / OR_G.poly_xt, OR_G.poly_alpha) / (OR_G.poly_alpha + 1.)) + 0.5 * M_PI;
    *th = thG + exp(OR_G.mks_smooth * (OR_G.startx[1] - X[1])) * (thJ - thG);
  }
}

/* src/coordinates.c:216-239 */
void or_gcov_ks(double r, double th, double gcov[4][4])
{
  double a = OR_G.a;
  double cth = cos(th);
  double sth = sin(th);
  double s2 = sth * sth;
  double rho2 = r * r + a * a * cth * cth;
  memset(gcov, 0, 16 * sizeof(double));
  gcov[0][0] = -1. + 2. * r / rho2;
  gcov[0][1] = 2. * r / rho2;
  gcov[0][3] = -2. * a * r * s2 / rho2;
  gcov[1][0] = gcov[0][1];
  gcov[1][1] = 1. + 2. * r / rho2;
  gcov[1][3] = -a * s2 * (1. + 2. * r / rho2);
  gcov[2][2] = rho2;
  gcov[3][0] = gcov[0][3];
  gcov[3][1] = gcov[1][3];
  gcov[3][3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2));
}

/* src/coordinates.c:241-261 */
void or_gcov_bl(double r, double th, double gcov[4][4])
{
  double a = OR_G.a;
  double sth = fabs(sin(th));
  double s2 = sth * sth;
  double cth = cos(th);
  double a2 = a * a;
  double r2 = r * r;
  double DD = 1. - 2. / r + a2 / r2;
  double mu = 1. + a2 * cth * cth / r2;
  memset(gcov, 0, 16 * sizeof(double));
  gcov[0][0] = -(1. - 2. / (r * mu));
  gcov[0][3] = -2. * a * s2 / (r * mu);
  gcov[3][0] = gcov[0][3];
  gcov[1][1] = mu / DD;
  gcov[2][2] = r2 * mu;
  gcov[3][3] = r2 * sth * sth * (1. + a2 / r2 + 2. * a2 * s2 / (r2 * r * mu));
}

/* src/coordinates.c:263-348 */
void or_set_dxdX(const double X[4], double dxdX[4][4])
{
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++) dxdX[mu][nu] = (mu == nu);
  dxdX[1][1] = exp(X[1]);
  double hs = OR_G.hslope, pn = OR_G.poly_norm, xt = OR_G.poly_xt, al = OR_G.poly_alpha;
  switch (OR_G.metric) {
  case OR_METRIC_MKS:
    dxdX[2][2] = M_PI + (1 - hs) * M_PI * cos(2. * M_PI * X[2]);
    break;
  case OR_METRIC_FMKS:
    dxdX[2][1] = -exp(OR_G.mks_smooth * (OR_G.startx[1] - X[1])) * OR_G.mks_smooth
        * (M_PI / 2. - M_PI * X[2]
           + pn * (2. * X[2] - 1.) * (1 + (pow((-1. + 2 * X[2]) / xt, al)) / (1 + al))
           - 1. / 2. * (1. - hs) * sin(2. * M_PI * X[2]));
    dxdX[2][2] = M_PI + (1. - hs) * M_PI * cos(2. * M_PI * X[2])
        + exp(OR_G.mks_smooth * (OR_G.startx[1] - X[1]))
          * (-M_PI + 2. * pn * (1. + pow((2. * X[2] - 1.) / xt, al) / (al + 1.))
             + (2. * al * pn * (2. * X[2] - 1.) * pow((2. * X[2] - 1.) / xt, al - 1.))
               / ((1. + al) * xt)
             - (1. - hs) * M_PI * cos(2. * M_PI * X[2]));
    break;
  case OR_METRIC_MINKOWSKI:
    dxdX[1][1] = 1.;
    break;
  case OR_METRIC_EMINKOWSKI:
    break;
  }
}

/* src/coordinates.c:175-213 */
void or_gcov(const double X[4], double gcov[4][4])
{
  double r, th;
  or_bl_coord(X, &r, &th);
  if (OR_G.metric == OR_METRIC_MINKOWSKI) {
    memset(gcov, 0, 16 * sizeof(double));
    gcov[0][0] = -1;
    gcov[1][1] = 1;
    gcov[2][2] = r * r;
    gcov[3][3] = r * r * sin(th) * sin(th);
    return;
  } else if (OR_G.metric == OR_METRIC_EMINKOWSKI) {
    memset(gcov, 0, 16 * sizeof(double));
    gcov[0][0] = -1;
    gcov[1][1] = r * r;
    gcov[2][2] = r * r;
    gcov[3][3] = r * r * sin(th) * sin(th);
    return;
  }
  double Gks[4][4], dxdX[4][4];
  or_gcov_ks(r, th, Gks);
  or_set_dxdX(X, dxdX);
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++) {
      gcov[mu][nu] = 0;
      for (int lam = 0; lam < 4; ++lam)
        for (int kap = 0; kap < 4; ++kap)
          gcov[mu][nu] += Gks[lam][kap] * dxdX[lam][mu] * dxdX[kap][nu];
    }
}

/* src/geometry.c:304-457, Crout LU with implicit pivoting */
static int lu_decompose(double A[4][4], int permute[4])
{
  const double absmin = 1.e-30;
  double row_norm[4];
  double absmax, maxtemp;
  int i, j, k, max_row = 0, n = 4;
  for (i = 0; i < n; i++) {
    absmax = 0.;
    for (j = 0; j < n; j++) {
      maxtemp = fabs(A[i][j]);
      if (maxtemp > absmax) absmax = maxtemp;
    }
    if (absmax == 0.) return 1;
    row_norm[i] = 1. / absmax;
  }
  for (j = 0; j < n; j++) {
    for (i = 0; i < j; i++)
      for (k = 0; k < i; k++) A[i][j] -= A[i][k] * A[k][j];
    absmax = 0.0;
    for (i = j; i < n; i++) {
      for (k = 0; k < j; k++) A[i][j] -= A[i][k] * A[k][j];
      maxtemp = fabs(A[i][j]) * row_norm[i];
      if (maxtemp >= absmax) {
        absmax = maxtemp;
        max_row = i;
      }
    }
    if (max_row != j) {
      if ((j == (n - 2)) && (A[j][j + 1] == 0.)) {
        max_row = j;
      } else {
        for (k = 0; k < n; k++) {
          maxtemp = A[j][k];
          A[j][k] = A[max_row][k];
          A[max_row][k] = maxtemp;
        }
        row_norm[max_row] = row_norm[j];
      }
    }
    permute[j] = max_row;
    if (A[j][j] == 0.) A[j][j] = absmin;
    if (j != (n - 1)) {
      maxtemp = 1. / A[j][j];
      for (i = (j + 1); i < n; i++) A[i][j] *= maxtemp;
    }
  }
  return 0;
}

/* src/geometry.c:471-504 */
static void lu_substitution(double A[4][4], double B[4], int permute[4])
{
  int i, j, n = 4;
  double tmpvar;
  for (i = 0; i < n; i++) {
    tmpvar = B[permute[i]];
    B[permute[i]] = B[i];
    for (j = (i - 1); j >= 0; j--) tmpvar -= A[i][j] * B[j];
    B[i] = tmpvar;
  }
  for (i = (n - 1); i >= 0; i--) {
    for (j = (i + 1); j < n; j++) B[i] -= A[i][j] * B[j];
    B[i] /= A[i][i];
  }
}

/* src/geometry.c:252-287 */
int or_invert_matrix(double Am[4][4], double Aminv[4][4])
{
  int permute[4];
  double dxm[4], Amtmp[4][4];
  memcpy(Amtmp, Am, sizeof(Amtmp));
  if (lu_decompose(Amtmp, permute) != 0) return 1;
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) dxm[j] = 0.;
    dxm[i] = 1.;
    lu_substitution(Amtmp, dxm, permute);
    for (int j = 0; j < 4; j++) Aminv[j][i] = dxm[j];
  }
  return 0;
}

/* src/geometry.c:54-64 */
int or_gcon(double gcov[4][4], double gcon[4][4]) { return or_invert_matrix(gcov, gcon); }

/* src/geometry.c:26-51 */
double or_gdet(double gcov[4][4])
{
  int permute[4];
  double tmp[4][4];
  memcpy(tmp, gcov, sizeof(tmp));
  if (lu_decompose(tmp, permute) != 0) return -1;
  double gdet = 1.;
  for (int i = 0; i < 4; i++) gdet *= tmp[i][i];
  return sqrt(fabs(gdet));
}

/* src/geometry.c:66-87 */
void or_lower(const double u[4], double g[4][4], double l[4])
{
  l[0] = g[0][0] * u[0] + g[0][1] * u[1] + g[0][2] * u[2] + g[0][3] * u[3];
  l[1] = g[1][0] * u[0] + g[1][1] * u[1] + g[1][2] * u[2] + g[1][3] * u[3];
  l[2] = g[2][0] * u[0] + g[2][1] * u[1] + g[2][2] * u[2] + g[2][3] * u[3];
  l[3] = g[3][0] * u[0] + g[3][1] * u[1] + g[3][2] * u[2] + g[3][3] * u[3];
}

/* src/geometry.c:91-144: central differences with DEL = 1e-7, then raise */
void or_connection(const double X[4], double conn[4][4][4])
{
  const double DEL = 1.e-7;
  double tmp[4][4][4], Xh[4], Xl[4], gcon[4][4], gcov[4][4], gh[4][4], gl[4][4];
  or_gcov(X, gcov);
  or_gcon(gcov, gcon);
  for (int k = 0; k < 4; k++) {
    for (int l = 0; l < 4; l++) Xh[l] = X[l];
    for (int l = 0; l < 4; l++) Xl[l] = X[l];
    Xh[k] += DEL;
    Xl[k] -= DEL;
    or_gcov(Xh, gh);
    or_gcov(Xl, gl);
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) conn[i][j][k] = (gh[i][j] - gl[i][j]) / (Xh[k] - Xl[k]);
  }
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++)
        tmp[i][j][k] = 0.5 * (conn[j][i][k] + conn[k][i][j] - conn[k][j][i]);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++) {
        conn[i][j][k] = 0.;
        for (int l = 0; l < 4; l++) conn[i][j][k] += gcon[i][l] * tmp[l][j][k];
      }
}

/* src/geometry.c:151-166 */
static void normalize(double v[4], double g[4][4])
{
  double norm = 0.;
  for (int k = 0; k < 4; k++)
    for (int l = 0; l < 4; l++) norm += v[k] * v[l] * g[k][l];
  norm = sqrt(fabs(norm));
  for (int k = 0; k < 4; k++) v[k] /= norm;
}

/* src/tetrads.c:95-115 */
static void project_out(double va[4], double vb[4], double g[4][4])
{
  double vb_sq = 0., adotb = 0.;
  for (int k = 0; k < 4; k++)
    for (int l = 0; l < 4; l++) vb_sq += vb[k] * vb[l] * g[k][l];
  for (int k = 0; k < 4; k++)
    for (int l = 0; l < 4; l++) adotb += va[k] * vb[l] * g[k][l];
  for (int k = 0; k < 4; k++) va[k] -= vb[k] * adotb / vb_sq;
}

/* src/tetrads.c:40-54 */
static void set_Econ_from_trial(double E[4], int defdir, const double trial[4])
{
  double norm = 0.;
  for (int k = 1; k < 4; k++) norm += fabs(trial[k]);
  for (int k = 0; k < 4; k++)
    E[k] = (norm <= OR_SMALL) ? (double)(k == defdir) : trial[k];
}

/* src/geometry.c:205-232 */
static int levi_civita(int i, int j, int k, int l)
{
  if (i == j || i == k || i == l || j == k || j == l || k == l) return 0;
  int idx[4] = {i, j, k, l}, n_perm = 0, do_sort = 1;
  while (do_sort) {
    int n_swap = 0;
    for (int n = 0; n < 3; n++)
      if (idx[n] > idx[n + 1]) {
        n_perm++;
        n_swap++;
        int v = idx[n];
        idx[n] = idx[n + 1];
        idx[n + 1] = v;
      }
    do_sort = n_swap;
  }
  return (n_perm % 2) ? -1 : 1;
}

/* src/tetrads.c:64-87 */
static int check_handedness(double Econ[4][4], double g[4][4], double *dot)
{
  double gd = or_gdet(g);
  if (gd < 0.) return 1;
  *dot = 0.;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int l = 0; l < 4; l++)
        for (int k = 0; k < 4; k++)
          *dot += gd * levi_civita(i, j, k, l) * Econ[0][i] * Econ[1][j] * Econ[2][k] * Econ[3][l];
  return 0;
}

/* src/model_tetrads.c:35-127 (eKS branch out of scope: use_eKS_internal = 0) */
int or_plasma_tetrad(const double U[4], const double K[4], const double B[4], double g[4][4],
                     double Econ[4][4], double Ecov[4][4])
{
  const double ones[4] = {1., 1., 1., 1.};
  set_Econ_from_trial(Econ[0], 0, U);
  set_Econ_from_trial(Econ[1], 3, ones);
  set_Econ_from_trial(Econ[2], 2, B);
  set_Econ_from_trial(Econ[3], 3, K);
  normalize(Econ[0], g);
  project_out(Econ[3], Econ[0], g);
  project_out(Econ[3], Econ[0], g);
  normalize(Econ[3], g);
  project_out(Econ[2], Econ[0], g);
  project_out(Econ[2], Econ[3], g);
  project_out(Econ[2], Econ[0], g);
  project_out(Econ[2], Econ[3], g);
  normalize(Econ[2], g);
  project_out(Econ[1], Econ[0], g);
  project_out(Econ[1], Econ[2], g);
  project_out(Econ[1], Econ[3], g);
  project_out(Econ[1], Econ[0], g);
  project_out(Econ[1], Econ[2], g);
  project_out(Econ[1], Econ[3], g);
  normalize(Econ[1], g);
  int oddflag = 0;
  double dot = 0.;
  if (check_handedness(Econ, g, &dot)) oddflag |= 16;
  if (fabs(fabs(dot) - 1.) > 1.e-10) oddflag |= 1;
  if (dot < 0.)
    for (int k = 0; k < 4; k++) Econ[1][k] *= -1.;
  for (int k = 0; k < 4; k++) or_lower(Econ[k], g, Ecov[k]);
  for (int l = 0; l < 4; l++) Ecov[0][l] *= -1.;
  return oddflag;
}

/* src/model_tetrads.c:143-193 */
int or_camera_tetrad(const double X[4], double Econ[4][4], double Ecov[4][4])
{
  double g[4][4], gc[4][4], Ucam[4], Kcon[4], trial[4];
  or_gcov(X, g);
  or_gcon(g, gc);
  trial[0] = -1.; trial[1] = 0.; trial[2] = 0.; trial[3] = 0.;
  or_lower(trial, gc, Ucam);
  trial[0] = 1.; trial[1] = 1.; trial[2] = 0.; trial[3] = 0.;
  or_lower(trial, gc, Kcon);
  trial[0] = 0.; trial[1] = 0.; trial[2] = 1.; trial[3] = 0.;
  return or_plasma_tetrad(Ucam, Kcon, trial, g, Econ, Ecov);
}

/* src/model_tetrads.c:209-249 */
int or_camera_tetrad_old(const double X[4], double Econ[4][4], double Ecov[4][4])
{
  double g[4][4], gc[4][4], Ucam[4] = {1., 0., 0., 0.}, Kcon[4], trial[4];
  or_gcov(X, g);
  or_gcon(g, gc);
  trial[0] = 1.; trial[1] = 1.; trial[2] = 0.; trial[3] = 0.;
  or_lower(trial, gc, Kcon);
  trial[0] = 0.; trial[1] = 0.; trial[2] = 1.; trial[3] = 0.;
  return or_plasma_tetrad(Ucam, Kcon, trial, g, Econ, Ecov);
}

/* src/geometry.c:235-240 */
static double theta_func(const double X[4])
{
  double r, th;
  or_bl_coord(X, &r, &th);
  return th;
}

/* src/coordinates.c:394-442 */
static double root_find(const double X[4])
{
  double th = X[2], tha, thb, thc;
  double Xa[4] = {0}, Xb[4] = {0}, Xc[4] = {0};
  Xa[1] = log(X[1]);
  Xa[3] = X[3];
  Xb[1] = Xa[1];
  Xb[3] = Xa[3];
  Xc[1] = Xa[1];
  Xc[3] = Xa[3];
  if (X[2] < M_PI / 2.) {
    Xa[2] = OR_G.cstartx[2];
    Xb[2] = (OR_G.cstopx[2] - OR_G.cstartx[2]) / 2 + OR_SMALL;
  } else {
    Xa[2] = (OR_G.cstopx[2] - OR_G.cstartx[2]) / 2 - OR_SMALL;
    Xb[2] = OR_G.cstopx[2];
  }
  double tol = 1.e-9;
  tha = theta_func(Xa);
  thb = theta_func(Xb);
  if (fabs(tha - th) < tol) return Xa[2];
  else if (fabs(thb - th) < tol) return Xb[2];
  for (int i = 0; i < 1000; i++) {
    Xc[2] = 0.5 * (Xa[2] + Xb[2]);
    thc = theta_func(Xc);
    if ((thc - th) * (thb - th) < 0.) Xa[2] = Xc[2];
    else Xb[2] = Xc[2];
    double err = theta_func(Xc) - th;
    if (fabs(err) < tol) break;
  }
  return Xc[2];
}

/* src/coordinates.c:375-387 */
void or_native_coord(double r, double th, double phi, double X[4])
{
  if (OR_G.metric == OR_METRIC_MINKOWSKI) {
    X[0] = 1; X[1] = r; X[2] = th / 180 * M_PI; X[3] = phi / 180 * M_PI;
  } else if (OR_G.metric == OR_METRIC_EMINKOWSKI) {
    X[0] = 1; X[1] = log(r); X[2] = th / 180 * M_PI; X[3] = phi / 180 * M_PI;
  } else {
    double x[4] = {0., r, th / 180. * M_PI, phi / 180. * M_PI};
    X[0] = 0.0;
    X[1] = log(r);
    X[2] = root_find(x);
    X[3] = phi / 180. * M_PI;
  }
}
