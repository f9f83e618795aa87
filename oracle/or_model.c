/*
 * Oracle (test infrastructure): fluid models and grid lookup.
 * Restates src/grid.c (ghost-aware trilinear interpolation),
 * model/iharm/model.c (fluid four-vectors, n_e, Theta_e, B, sigma, the
 * one-time dump derivation and ghost-zone boundary conditions),
 * model/sphere/model.c and model/analytic/model.c.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

extern or_geom_t OR_G;
or_model_t OR_M;

void or_set_model(const or_model_t *m) { OR_M = *m; }

/* ---------------------------------------------------------------- grid */

static inline size_t gidx(int i, int j, int k)
{
  return ((size_t)i * (OR_G.N2 + 2) + j) * (size_t)(OR_G.N3 + 2) + k;
}

/* src/grid.c:216-283 (plain native coordinates: no eKS, no simcoords) */
static void Xtoijk_ghost(const double X[4], int *i, int *j, int *k, double del[4])
{
  double phi = fmod(X[3], OR_G.cstopx[3]);
  if (phi < 0.0) phi = OR_G.cstopx[3] + phi;
  *i = (int)((X[1] - OR_G.startx[1]) / OR_G.dx[1] - 0.5 + 1000) - 1000;
  *j = (int)((X[2] - OR_G.startx[2]) / OR_G.dx[2] - 0.5 + 1000) - 1000;
  *k = (int)((phi - OR_G.startx[3]) / OR_G.dx[3] - 0.5 + 1000) - 1000;
  if (*i < -1) *i = -1;
  if (*j < -1) *j = -1;
  if (*k < -1) *k = -1;
  if (*i >= OR_G.N1) *i = OR_G.N1 - 1;
  if (*j >= OR_G.N2) *j = OR_G.N2 - 1;
  if (*k >= OR_G.N3) *k = OR_G.N3 - 1;
  del[0] = 0.;
  del[1] = (X[1] - ((*i + 0.5) * OR_G.dx[1] + OR_G.startx[1])) / OR_G.dx[1];
  del[2] = (X[2] - ((*j + 0.5) * OR_G.dx[2] + OR_G.startx[2])) / OR_G.dx[2];
  del[3] = (phi - ((*k + 0.5) * OR_G.dx[3] + OR_G.startx[3])) / OR_G.dx[3];
  for (int n = 0; n < 4; ++n) {
    if (del[n] < 0.) del[n] = 0.;
    if (del[n] >= 1.) del[n] = 1.;
  }
  *i += 1;
  *j += 1;
  *k += 1;
}

/* src/grid.c:15-44 */
static double interp_scalar(const double X[4], const double *var)
{
  double del[4];
  int i, j, k;
  Xtoijk_ghost(X, &i, &j, &k, del);
  int ip1 = i + 1, jp1 = j + 1, kp1 = k + 1;
  double b1 = 1. - del[1];
  double b2 = 1. - del[2];
  double interp = var[gidx(i, j, k)] * b1 * b2 + var[gidx(i, jp1, k)] * b1 * del[2]
      + var[gidx(ip1, j, k)] * del[1] * b2 + var[gidx(ip1, jp1, k)] * del[1] * del[2];
  interp = (1. - del[3]) * interp
      + del[3] * (var[gidx(i, j, kp1)] * b1 * b2 + var[gidx(i, jp1, kp1)] * del[2] * b1
                  + var[gidx(ip1, j, kp1)] * del[1] * b2 + var[gidx(ip1, jp1, kp1)] * del[1] * del[2]);
  return interp;
}

/* src/grid.c:285-334 */
static int X_in_domain(const double X[4])
{
  if (X[1] < OR_G.startx[1] || X[1] >= OR_G.stopx[1] || X[2] < OR_G.startx[2]
      || X[2] >= OR_G.stopx[2])
    return 0;
  return 1;
}

static inline const double *field(int f)
{
  size_t n = (size_t)(OR_G.N1 + 2) * (OR_G.N2 + 2) * (OR_G.N3 + 2);
  return OR_M.fields + (size_t)f * n;
}

/* ---------------------------------------------------------------- iharm */

/* model/iharm/model.c:420-493 */
static void iharm_fourv(const double X[4], double Ucon[4], double Ucov[4], double Bcon[4],
                        double Bcov[4])
{
  double gcov[4][4], gcon[4][4];
  or_gcov(X, gcov);
  or_gcon(gcov, gcon);
  if (X_in_domain(X) == 0) {
    Ucov[0] = -1. / sqrt(-gcon[0][0]);
    Ucov[1] = 0.;
    Ucov[2] = 0.;
    Ucov[3] = 0.;
    Ucon[0] = 0.;
    Ucon[1] = 0.;
    Ucon[2] = 0.;
    Ucon[3] = 0.;
    for (int mu = 0; mu < 4; ++mu) {
      Ucon[0] += Ucov[mu] * gcon[0][mu];
      Ucon[1] += Ucov[mu] * gcon[1][mu];
      Ucon[2] += Ucov[mu] * gcon[2][mu];
      Ucon[3] += Ucov[mu] * gcon[3][mu];
      Bcon[mu] = 0.;
      Bcov[mu] = 0.;
    }
    return;
  }
  double Vcon[4];
  Vcon[1] = interp_scalar(X, field(0));
  Vcon[2] = interp_scalar(X, field(1));
  Vcon[3] = interp_scalar(X, field(2));
  double VdotV = 0.;
  for (int i = 1; i < 4; i++)
    for (int j = 1; j < 4; j++) VdotV += gcov[i][j] * Vcon[i] * Vcon[j];
  double Vfac = sqrt(-1. / gcon[0][0] * (1. + fabs(VdotV)));
  Ucon[0] = -Vfac * gcon[0][0];
  for (int i = 1; i < 4; i++) Ucon[i] = Vcon[i] - Vfac * gcon[0][i];
  or_lower(Ucon, gcov, Ucov);
  double Bcon1 = interp_scalar(X, field(3));
  double Bcon2 = interp_scalar(X, field(4));
  double Bcon3 = interp_scalar(X, field(5));
  Bcon[0] = Bcon1 * Ucov[1] + Bcon2 * Ucov[2] + Bcon3 * Ucov[3];
  Bcon[1] = (Bcon1 + Ucon[1] * Bcon[0]) / Ucon[0];
  Bcon[2] = (Bcon2 + Ucon[2] * Bcon[0]) / Ucon[0];
  Bcon[3] = (Bcon3 + Ucon[3] * Bcon[0]) / Ucon[0];
  or_lower(Bcon, gcov, Bcov);
}

/* model/iharm/model.c:567-575 */
static double sigma_smoothfac(double sigma)
{
  double sigma_above = OR_M.sigma_cut;
  if (OR_M.sigma_cut_high > 0) sigma_above = OR_M.sigma_cut_high;
  if (sigma < OR_M.sigma_cut) return 1;
  if (sigma >= sigma_above) return 0;
  double dsig = sigma_above - OR_M.sigma_cut;
  return cos(M_PI / 2. / dsig * (sigma - OR_M.sigma_cut));
}

/* ---------------------------------------------------------------- analytic */

/* model/analytic/model.c:233-289 (bl_to_ks: src/coordinates.c:99-115,
 * vec_from_ks: :350-370) */
static void analytic_fourv(const double X[4], double Ucon[4], double Ucov[4], double Bcon[4],
                           double Bcov[4])
{
  double r, th;
  or_bl_coord(X, &r, &th);
  double R = r * sin(th);
  double l = (OR_M.an_l0 / (1 + R)) * pow(R, 1 + 0.5);
  double bl_gcov[4][4], bl_gcon[4][4], gcov[4][4], gcon[4][4];
  or_gcov_bl(r, th, bl_gcov);
  or_gcon(bl_gcov, bl_gcon);
  or_gcov(X, gcov);
  or_gcon(gcov, gcon);
  double bl_Ucov[4];
  double ubar = sqrt(-1. / (bl_gcon[0][0] - 2. * bl_gcon[0][3] * l + bl_gcon[3][3] * l * l));
  bl_Ucov[0] = -ubar;
  bl_Ucov[1] = 0.;
  bl_Ucov[2] = 0.;
  bl_Ucov[3] = l * ubar;
  double bl_Ucon[4];
  or_lower(bl_Ucov, bl_gcon, bl_Ucon);
  /* bl_to_ks */
  double trans[4][4];
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++) trans[mu][nu] = (mu == nu);
  double a = OR_G.a;
  trans[0][1] = 2. * r / (r * r - 2. * r + a * a);
  trans[3][1] = a / (r * r - 2. * r + a * a);
  double ks_Ucon[4] = {0., 0., 0., 0.};
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++) ks_Ucon[mu] += trans[mu][nu] * bl_Ucon[nu];
  /* vec_from_ks */
  double dxdX[4][4], dXdx[4][4];
  or_set_dxdX(X, dxdX);
  or_invert_matrix(dxdX, dXdx);
  for (int mu = 0; mu < 4; mu++) Ucon[mu] = 0.;
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++) Ucon[mu] += dXdx[mu][nu] * ks_Ucon[nu];
  or_lower(Ucon, gcov, Ucov);
  Bcon[0] = 0;
  Bcon[1] = 0;
  Bcon[2] = 1;
  Bcon[3] = 0;
  or_lower(Bcon, gcov, Bcov);
}

/* ---------------------------------------------------------------- dispatch */

/* model/iharm/model.c:2093-2098, model/sphere/model.c:213-218,
 * model/analytic/model.c:315-322 */
int or_radiating_region(const double X[4])
{
  double r, th;
  or_bl_coord(X, &r, &th);
  if (OR_M.kind == OR_MODEL_IHARM)
    return (r > OR_G.rmin_geo && r < OR_G.rmax_geo && th > OR_M.th_beg
            && th < (M_PI - OR_M.th_beg));
  if (OR_M.kind == OR_MODEL_SPHERE) return (r < OR_M.sph_R0) ? 1 : 0;
  return r > OR_G.Rh + 0.0001 && r > OR_G.rmin_geo && r < 1000.0;
}

/* model/sphere/model.c:126-148 */
static void sphere_fourv(const double X[4], double Ucon[4], double Ucov[4], double Bcon[4],
                         double Bcov[4])
{
  double r, h, gcov[4][4];
  or_bl_coord(X, &r, &h);
  or_gcov(X, gcov);
  Ucon[0] = 1.;
  Ucon[1] = 0.;
  Ucon[2] = 0.;
  Ucon[3] = 0.;
  Bcon[0] = 0.;
  Bcon[1] = OR_M.sph_B0 * cos(h) / r;
  Bcon[2] = -OR_M.sph_B0 * sin(h) / (r + 1.e-8);
  Bcon[3] = 0.;
  or_lower(Ucon, gcov, Ucov);
  or_lower(Bcon, gcov, Bcov);
}

void or_fourv(const double X[4], const double K[4], double Ucon[4], double Ucov[4],
              double Bcon[4], double Bcov[4])
{
  (void)K;
  if (OR_M.kind == OR_MODEL_IHARM) iharm_fourv(X, Ucon, Ucov, Bcon, Bcov);
  else if (OR_M.kind == OR_MODEL_SPHERE) sphere_fourv(X, Ucon, Ucov, Bcon, Bcov);
  else analytic_fourv(X, Ucon, Ucov, Bcon, Bcov);
}

/* model/iharm/model.c:545-553 */
double or_sigma(const double X[4])
{
  if (OR_M.kind == OR_MODEL_SPHERE) return or_radiating_region(X) ? OR_M.sph_sigma : 0.;
  if (OR_M.kind != OR_MODEL_IHARM) return 0.;
  if (X_in_domain(X) == 0) return 0.;
  return interp_scalar(X, field(9));
}

/* model/iharm/model.c:577-593; model/sphere/model.c:205-211;
 * model/analytic/model.c:175-183 */
double or_ne(const double X[4])
{
  if (OR_M.kind == OR_MODEL_IHARM) {
    if (X_in_domain(X) == 0) return 0.;
    double sigma = or_sigma(X);
    if (sigma > OR_M.sigma_cut) return 0.;
    double sf = sigma_smoothfac(sigma);
    return interp_scalar(X, field(6)) * sf;
  }
  if (OR_M.kind == OR_MODEL_SPHERE) return or_radiating_region(X) ? OR_M.sph_Ne0 : 0.;
  double r, th;
  or_bl_coord(X, &r, &th);
  double n_exp = 1. / 2 * (pow(r / 10, 2) + pow(OR_M.an_height * cos(th), 2));
  return (n_exp < 200) ? OR_M.an_RHO_unit * exp(-n_exp) : 0;
}

/* model/iharm/model.c:510-531; model/sphere/model.c:162-169 */
double or_thetae(const double X[4])
{
  if (OR_M.kind == OR_MODEL_IHARM) {
    if (X_in_domain(X) == 0) return 0.;
    return interp_scalar(X, field(7));
  }
  if (OR_M.kind == OR_MODEL_SPHERE) return or_radiating_region(X) ? OR_M.sph_Thetae0 : 0.;
  return 0.;
}

/* model/iharm/model.c:534-543; model/sphere/model.c:172-179;
 * model/analytic/model.c:297-304 */
double or_b(const double X[4])
{
  if (OR_M.kind == OR_MODEL_IHARM) {
    if (X_in_domain(X) == 0) return 0.;
    return interp_scalar(X, field(8));
  }
  if (OR_M.kind == OR_MODEL_SPHERE) return or_radiating_region(X) ? OR_M.sph_B0 : 0.;
  double Ucon[4], Ucov[4], Bcon[4], Bcov[4], K[4] = {0};
  or_fourv(X, K, Ucon, Ucov, Bcon, Bcov);
  return sqrt(Bcon[0] * Bcov[0] + Bcon[1] * Bcov[1] + Bcon[2] * Bcov[2] + Bcon[3] * Bcov[3])
      * OR_M.an_B_unit;
}

/* get_model_primitives (model/iharm/model.c:498-508) and get_model_beta
 * (:555-565) of the iharm model: rho, u interpolated from rub[0], rub[1],
 * U1..B3 from the model fields, beta from rub[2] (rub: 3 ghosted arrays, or
 * NULL, which leaves rho, u and beta 0).  The primitives stay 0 outside the
 * domain and for other models; beta is 0 there too. */
void or_primitives(const double X[4], const double *rub, double p8[8], double *beta)
{
  for (int n = 0; n < 8; n++) p8[n] = 0.;
  *beta = 0.;
  if (OR_M.kind == OR_MODEL_SPHERE && or_radiating_region(X)) *beta = OR_M.sph_beta;
  if (OR_M.kind != OR_MODEL_IHARM || X_in_domain(X) == 0) return;
  const size_t nc = (size_t)(OR_G.N1 + 2) * (OR_G.N2 + 2) * (OR_G.N3 + 2);
  for (int n = 0; n < 6; n++) p8[2 + n] = interp_scalar(X, field(n));
  if (rub) {
    p8[0] = interp_scalar(X, rub);
    p8[1] = interp_scalar(X, rub + nc);
    *beta = interp_scalar(X, rub + 2 * nc);
  }
}

/* ---------------------------------------------------------------- setup */

/* model/sphere/model.c:58-96 */
void or_sphere_setup(double R0, double tau0, double Thetae0, double beta0, double gam,
                     double tp_over_te, double MBH_solar, or_model_t *m)
{
  memset(m, 0, sizeof(*m));
  m->kind = OR_MODEL_SPHERE;
  double MBH = MBH_solar * OR_MSUN;
  m->L_unit = OR_GNEWT * MBH / (OR_CL * OR_CL);
  m->sph_R0 = R0;
  m->sph_Ne0 = tau0 / OR_SIGMA_THOMSON / R0 / m->L_unit;
  double THETAE_UNIT = OR_MP / OR_ME * (gam - 1.) / (1. + tp_over_te);
  m->sph_B0 = OR_CL * sqrt(8 * M_PI * (gam - 1.) * (OR_MP + OR_ME) / beta0)
      * sqrt(m->sph_Ne0 * Thetae0) / sqrt(THETAE_UNIT);
  m->sph_Thetae0 = Thetae0;
  m->sph_sigma = Thetae0 * tp_over_te / beta0;
  m->sph_beta = beta0;
}

/* model/analytic/model.c:70-150 */
void or_analytic_setup(int model, double MBH_solar, double rho_unit, double freqcgs,
                       or_model_t *m, or_geom_t *g)
{
  memset(m, 0, sizeof(*m));
  memset(g, 0, sizeof(*g));
  m->kind = OR_MODEL_ANALYTIC;
  m->an_model = model;
  g->metric = OR_METRIC_MKS;
  g->hslope = 1.0;
  double A = 0, alpha = 0, height = 0, l0 = 0, a = 0;
  if (model == 1) { A = 0; alpha = -3; height = 0; l0 = 0; a = 0.9; }
  else if (model == 2) { A = 0; alpha = -2; height = 0; l0 = 1; a = 0; }
  else if (model == 3) { A = 0; alpha = 0; height = 10. / 3; l0 = 1; a = 0.9; }
  else if (model == 4) { A = 1.e5; alpha = 0; height = 10. / 3; l0 = 1; a = 0.9; }
  else if (model == 5) { A = 1.e6; alpha = 0; height = 100. / 3; l0 = 1; a = 0.9; }
  m->an_A = A;
  m->an_alpha = alpha;
  m->an_height = height;
  m->an_l0 = l0;
  m->an_freqcgs = freqcgs;
  g->a = a;
  double MBH = MBH_solar * OR_MSUN;
  m->L_unit = OR_GNEWT * MBH / (OR_CL * OR_CL);
  m->an_RHO_unit = (rho_unit > 0) ? rho_unit : 3.e-18;
  m->an_B_unit = OR_CL * sqrt(4. * M_PI * m->an_RHO_unit);
  g->Rh = 1 + sqrt(1. - a * a);
  g->Rin = g->Rh;
  g->Rout = 1000.0;
  g->rmin_geo = 1.;
  g->rmax_geo = 10000.;
  g->startx[1] = log(g->Rin);
  g->stopx[1] = log(g->Rout);
  g->stopx[2] = 1.0;
  g->stopx[3] = 2 * M_PI;
  for (int mu = 0; mu < 4; mu++) {
    g->cstartx[mu] = g->startx[mu];
    g->cstopx[mu] = g->stopx[mu];
  }
}

/* ---------------------------------------------------------------- derivation */

/* One-time derivation of the ghosted fluid fields from a primitive dump:
 * model/iharm/model.c:595-608 (units), :1954-2010 (b from the primitives
 * at zone centers), :1314-1370 (ghost zones) and :610-717 (n_e, Theta_e with
 * the R_high/R_low model ELECTRONS = 2, sigma, beta).
 * prims: [N1][N2][N3][nprim] with rho, u, U1..U3, B1..B3 first.
 * out11: 11 ghosted arrays U1 U2 U3 B1 B2 B3 ne thetae b sigma beta. */
void or_iharm_derive(const double *prims, int nprim, const or_units_t *u, double *out11)
{
  or_iharm_derive_ru(prims, nprim, u, out11, NULL);
}

/* The same, also returning the ghosted rho and u arrays (out_ru: 2 x ncell,
 * may be NULL), which get_model_primitives interpolates for the trace output
 * (model/iharm/model.c:498-508). */
void or_iharm_derive_ru(const double *prims, int nprim, const or_units_t *u, double *out11,
                        double *out_ru)
{
  const int N1 = OR_G.N1, N2 = OR_G.N2, N3 = OR_G.N3;
  const size_t nc = (size_t)(N1 + 2) * (N2 + 2) * (N3 + 2);
  double MBH = u->MBH_solar * OR_MSUN;
  double L_unit = OR_GNEWT * MBH / (OR_CL * OR_CL);
  double RHO_unit = u->M_unit / pow(L_unit, 3);
  double B_unit = OR_CL * sqrt(4. * M_PI * RHO_unit);
  double *p[8];
  double *pbuf = calloc(8 * nc, sizeof(double));
  for (int l = 0; l < 8; l++) p[l] = pbuf + l * nc;
  double *b = out11 + 8 * nc;
  memset(out11, 0, 11 * nc * sizeof(double));
  for (int i = 0; i < N1; i++)
    for (int j = 0; j < N2; j++)
      for (int k = 0; k < N3; k++)
        for (int l = 0; l < 8; l++)
          p[l][gidx(i + 1, j + 1, k + 1)] = prims[(((size_t)i * N2 + j) * N3 + k) * nprim + l];

#pragma omp parallel for collapse(2)
  for (int i = 1; i < N1 + 1; i++)
    for (int j = 1; j < N2 + 1; j++) {
      double X[4] = {0.}, gcov[4][4], gcon[4][4];
      X[1] = OR_G.startx[1] + (i - 1 + 0.5) * OR_G.dx[1];
      X[2] = OR_G.startx[2] + (j - 1 + 0.5) * OR_G.dx[2];
      X[3] = OR_G.startx[3] + (0 + 0.5) * OR_G.dx[3];
      or_gcov(X, gcov);
      or_gcon(gcov, gcon);
      for (int k = 1; k < N3 + 1; k++) {
        size_t c = gidx(i, j, k);
        double UdotU = 0.;
        for (int l = 1; l < 4; l++)
          for (int m = 1; m < 4; m++) UdotU += gcov[l][m] * p[1 + l][c] * p[1 + m][c];
        double ufac = sqrt(-1. / gcon[0][0] * (1 + fabs(UdotU)));
        double ucon[4] = {0.};
        ucon[0] = -ufac * gcon[0][0];
        for (int l = 1; l < 4; l++) ucon[l] = p[1 + l][c] - ufac * gcon[0][l];
        double ucov[4];
        or_lower(ucon, gcov, ucov);
        double udotB = 0.;
        for (int l = 1; l < 4; l++) udotB += ucov[l] * p[4 + l][c];
        double bcon[4], bcov[4];
        bcon[0] = udotB;
        for (int l = 1; l < 4; l++) bcon[l] = (p[4 + l][c] + ucon[l] * udotB) / ucon[0];
        or_lower(bcon, gcov, bcov);
        double bsq = 0.;
        for (int l = 0; l < 4; ++l) bsq += bcon[l] * bcov[l];
        b[c] = sqrt(bsq) * B_unit;
      }
    }

  /* ghost zones, model/iharm/model.c:1314-1370 */
  for (int j = 1; j < N2 + 1; ++j)
    for (int k = 1; k < N3 + 1; ++k) {
      for (int l = 0; l < 8; ++l) {
        p[l][gidx(0, j, k)] = p[l][gidx(1, j, k)];
        p[l][gidx(N1 + 1, j, k)] = p[l][gidx(N1, j, k)];
      }
      b[gidx(0, j, k)] = b[gidx(1, j, k)];
      b[gidx(N1 + 1, j, k)] = b[gidx(N1, j, k)];
    }
  for (int i = 0; i < N1 + 2; ++i)
    for (int k = 1; k < N3 + 1; ++k) {
      if (N3 % 2 == 0) {
        int kflip = ((k - 1) + (N3 / 2)) % N3 + 1;
        for (int l = 0; l < 8; ++l) {
          p[l][gidx(i, 0, k)] = p[l][gidx(i, 1, kflip)];
          p[l][gidx(i, N2 + 1, k)] = p[l][gidx(i, N2, kflip)];
        }
        b[gidx(i, 0, k)] = b[gidx(i, 1, kflip)];
        b[gidx(i, N2 + 1, k)] = b[gidx(i, N2, kflip)];
      } else {
        int kflip1 = (k + (N3 / 2)) % N3;
        int kflip2 = (k + (N3 / 2) + 1) % N3;
        for (int l = 0; l < 8; ++l) {
          p[l][gidx(i, 0, k)] = (p[l][gidx(i, 1, kflip1)] + p[l][gidx(i, 1, kflip2)]) / 2.;
          p[l][gidx(i, N2 + 1, k)] = (p[l][gidx(i, N2, kflip1)] + p[l][gidx(i, N2, kflip2)]) / 2.;
        }
        b[gidx(i, 0, k)] = (b[gidx(i, 1, kflip1)] + b[gidx(i, 1, kflip2)]) / 2.;
        b[gidx(i, N2 + 1, k)] = (b[gidx(i, N2, kflip1)] + b[gidx(i, N2, kflip2)]) / 2.;
      }
    }
  for (int i = 0; i < N1 + 2; ++i)
    for (int j = 0; j < N2 + 2; ++j) {
      for (int l = 0; l < 8; ++l) {
        p[l][gidx(i, j, 0)] = p[l][gidx(i, j, N3)];
        p[l][gidx(i, j, N3 + 1)] = p[l][gidx(i, j, 1)];
      }
      b[gidx(i, j, 0)] = b[gidx(i, j, N3)];
      b[gidx(i, j, N3 + 1)] = b[gidx(i, j, 1)];
    }

  double *ne = out11 + 6 * nc, *thetae = out11 + 7 * nc, *sigma = out11 + 9 * nc,
         *beta = out11 + 10 * nc;
  double gam = u->gam, game = u->game, gamp = u->gamp;
#pragma omp parallel for
  for (size_t c = 0; c < nc; c++) {
    double rho = p[0][c], uu = p[1][c];
    ne[c] = rho * RHO_unit / (OR_MP + OR_ME) * 1.;
    double bsq = b[c] / B_unit;
    bsq = bsq * bsq;
    double sigma_m = bsq / rho;
    double beta_m = uu * (gam - 1.) / 0.5 / bsq;
    double betasq = beta_m * beta_m / u->beta_crit / u->beta_crit;
    double trat = u->trat_large * betasq / (1. + betasq) + u->trat_small / (1. + betasq);
    double lcl_Thetae_u = (OR_MP / OR_ME) * (game - 1.) * (gamp - 1.)
        / ((gamp - 1.) + (game - 1.) * trat);
    thetae[c] = lcl_Thetae_u * uu / rho;
    thetae[c] = fmax(thetae[c], 1.e-3);
    sigma[c] = fmax(sigma_m, OR_SMALL);
    beta[c] = fmax(beta_m, OR_SMALL);
  }
  for (int l = 0; l < 6; l++) memcpy(out11 + l * nc, p[2 + l], nc * sizeof(double));
  if (out_ru) {
    memcpy(out_ru, p[0], nc * sizeof(double));
    memcpy(out_ru + nc, p[1], nc * sizeof(double));
  }
  free(pbuf);
}
