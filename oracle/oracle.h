/*
 * CPU oracle for the ipole per-pixel polarized ray pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (AFD-Illinois/ipole) used as the checker for the GPU library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product path (ipole_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  It keeps the
 * reference's as-written arithmetic (finite-difference connection with
 * DEL = 1e-7 and a Crout LU inverse, full rank-2 complex tetrad transforms,
 * per-call model lookups), so that it can be compared with a build of the
 * reference's own translation units (oracle/_ref) step by step.
 *
 * Parity pins (see tests/): metric, connection, geodesic, tetrad, grid
 * lookup, dump derivation, push_polar, approximate_solve and project_N are
 * compared with oracle/_ref; K_n against scipy/mpmath; the Stokes step
 * against the matrix exponential of the Mueller generator; the thermal
 * fits against the cold-plasma limits and Kirchhoff's law; whole images
 * against the Gold et al. (2020) fluxes held in tests/analytic of the
 * reference and its isothermal-sphere tables.
 */
#ifndef IPOLE_ORACLE_H
#define IPOLE_ORACLE_H

#include <complex.h>
#include <stddef.h>

#define OR_NDIM 4
#define OR_NIMG 5

/* coordinate systems: same codes as src/coordinates.h:450-464 */
#define OR_METRIC_MKS 0
#define OR_METRIC_FMKS 2
#define OR_METRIC_MINKOWSKI 4
#define OR_METRIC_EMINKOWSKI 5

#define OR_MODEL_IHARM 1
#define OR_MODEL_SPHERE 2
#define OR_MODEL_ANALYTIC 3

/* constants: src/constants.h:3-62 */
#define OR_EE (4.80320680e-10)
#define OR_CL (2.99792458e10)
#define OR_ME (9.1093826e-28)
#define OR_MP (1.67262171e-24)
#define OR_HPL (6.6260693e-27)
#define OR_GNEWT (6.6742e-8)
#define OR_SIGMA_THOMSON (0.665245873e-24)
#define OR_JY (1.e-23)
#define OR_MUAS_PER_RAD (2.06265e11)
#define OR_PC (3.085678e18)
#define OR_MSUN (1.989e33)
#define OR_SMALL 1.e-40

typedef struct {
  int metric;
  double a, hslope, poly_xt, poly_alpha, mks_smooth, poly_norm;
  double startx[4], stopx[4], dx[4], cstartx[4], cstopx[4];
  double Rin, Rout, Rh, rmin_geo, rmax_geo;
  int N1, N2, N3;
} or_geom_t;

/* Fluid model.  iharm fields are ghosted (N1+2)(N2+2)(N3+2) arrays, x3
 * fastest, in the order U1 U2 U3 B1 B2 B3 ne thetae b sigma. */
typedef struct {
  int kind;
  double L_unit;                 /* cm per r_g */
  /* iharm (model/iharm/model.c) */
  const double *fields;          /* 10 consecutive ghosted arrays */
  double th_beg, sigma_cut, sigma_cut_high;
  /* sphere (model/sphere/model.c) */
  double sph_R0, sph_Ne0, sph_Thetae0, sph_B0;
  /* analytic (model/analytic/model.c) */
  double sph_sigma, sph_beta;    /* model/sphere/model.c:172-188 */
  int an_model;
  double an_A, an_alpha, an_height, an_l0, an_freqcgs, an_RHO_unit, an_B_unit;
} or_model_t;

/* Per-image parameters (src/par.h:14-73 subset on the fast-light path) */
typedef struct {
  int nx, ny;
  double rcam, thetacam, phicam, rotcam;     /* degrees, as read from par */
  double fovx_dsource, fovy_dsource, dx, dy; /* muas / r_g */
  double dsource;                             /* pc */
  double freqcgs;
  double eps;
  int maxnstep;
  int old_centering, only_unpolarized, qu_conv, stokes_floors;
  int target_nturns, subring_dtheta, counterjet, emission_type;
  double xoff, yoff;
} or_params_t;

/* derived camera quantities, src/main.c:126-174 */
typedef struct {
  double Xcam[4];
  double freq, fovx, fovy, DX, DY, scale, rotcam_rad, dsource_cm;
  int bad_pixel_i, bad_pixel_j;
} or_camera_t;

/* electron/unit parameters for the dump derivation (model/iharm/model.c) */
typedef struct {
  double M_unit, MBH_solar;
  double trat_small, trat_large, beta_crit;
  double gam, game, gamp;
} or_units_t;

typedef struct {
  int nturns;
  double dl;
  double X[4], Kcon[4], Xhalf[4], Kconhalf[4];
} or_traj_t;

/* state */
void or_set_geom(const or_geom_t *g);
void or_get_geom(or_geom_t *g);
void or_set_model(const or_model_t *m);

/* coordinates / metric (src/coordinates.c, src/geometry.c) */
void or_bl_coord(const double X[4], double *r, double *th);
void or_gcov(const double X[4], double g[4][4]);
void or_gcov_ks(double r, double th, double g[4][4]);
void or_gcov_bl(double r, double th, double g[4][4]);
void or_set_dxdX(const double X[4], double dxdX[4][4]);
int or_gcon(double gcov[4][4], double gcon[4][4]);
int or_invert_matrix(double Am[4][4], double Aminv[4][4]);
double or_gdet(double gcov[4][4]);
void or_connection(const double X[4], double conn[4][4][4]);
void or_native_coord(double r, double th, double phi, double X[4]);
void or_lower(const double u[4], double g[4][4], double l[4]);

/* geodesics (src/geodesics.c, src/model_geodesics.c) */
void or_push_photon(double X[4], double K[4], double dl, double Xh[4], double Kh[4]);
double or_stepsize(const double X[4], const double K[4], double eps);
int or_stop_backward(const double X[4], const double K[4]);
int or_trace(const double Xi[4], const double Ki[4], or_traj_t *traj, double eps,
             int step_max, const double Xcam[4], int subring_dtheta);
void or_init_XK(long i, long j, const or_params_t *p, const or_camera_t *cam,
                double X[4], double K[4]);

/* tetrads (src/tetrads.c, src/model_tetrads.c) */
int or_plasma_tetrad(const double U[4], const double K[4], const double B[4],
                     double g[4][4], double Econ[4][4], double Ecov[4][4]);
int or_camera_tetrad(const double X[4], double Econ[4][4], double Ecov[4][4]);
int or_camera_tetrad_old(const double X[4], double Econ[4][4], double Ecov[4][4]);

/* model */
int or_radiating_region(const double X[4]);
void or_fourv(const double X[4], const double K[4], double Ucon[4], double Ucov[4],
              double Bcon[4], double Bcov[4]);
double or_ne(const double X[4]);
double or_thetae(const double X[4]);
double or_b(const double X[4]);
double or_sigma(const double X[4]);
void or_iharm_derive(const double *prims, int nprim, const or_units_t *u,
                     double *out11);
void or_iharm_derive_ru(const double *prims, int nprim, const or_units_t *u,
                        double *out11, double *out_ru);
void or_primitives(const double X[4], const double *rub, double p8[8], double *beta);
void or_sphere_setup(double R0, double tau0, double Thetae0, double beta0, double gam,
                     double tp_over_te, double MBH_solar, or_model_t *m);
void or_analytic_setup(int model, double MBH_solar, double rho_unit, double freqcgs,
                       or_model_t *m, or_geom_t *g);

/* radiation (src/model_radiation.c, src/radiation.c, src/symphony) */
double or_bessel_kn(int n, double x);
void or_jar(int dist, int pol, const double X[4], const double K[4], double out[11]);

/* transfer (src/ipolarray.c) */
double or_approximate_solve(double Ii, double ji, double ki, double jf, double kf,
                            double dl, double *tau);
void or_push_polar(const double Xm[4], const double Km[4], double complex Ni[4][4],
                   double complex Nm[4][4], double complex Nf[4][4], double dlam);
void or_stokes_step(const double c11[11], double dlam, double S[4]);
int or_evolve_N(const double Xf[4], const double Kf[4], double dlam,
                double complex N[4][4], double *tauF, int zero_emission,
                const or_params_t *p);
void or_project_N(const double X[4], double complex N[4][4], double *SI, double *SQ,
                  double *SU, double *SV, double rotcam);
int or_integrate_emission(or_traj_t *traj, int nsteps, double *Intensity, double *tau,
                          double *tauF, double complex N[4][4], const or_params_t *p);

/* trace output (src/io.c:279-543): per-step record of OR_TR_NF doubles */
#define OR_TR_NF 126
#define OR_TR_PRIMS 0
#define OR_TR_B 8
#define OR_TR_NE 9
#define OR_TR_THETAE 10
#define OR_TR_NU 11
#define OR_TR_BK_ANGLE 12
#define OR_TR_SIGMA 13
#define OR_TR_BETA 14
#define OR_TR_DL 15
#define OR_TR_R 16
#define OR_TR_TH 17
#define OR_TR_PHI 18
#define OR_TR_J_UNPOL 19
#define OR_TR_K_UNPOL 20
#define OR_TR_I_UNPOL 21
#define OR_TR_X 22
#define OR_TR_KCON 26
#define OR_TR_J_INV 30
#define OR_TR_ALPHA_INV 34
#define OR_TR_RHO_INV 38
#define OR_TR_STOKES 42
#define OR_TR_E2C 46
#define OR_TR_E2C_COV 50
#define OR_TR_E2T 54
#define OR_TR_E2T_COV 58
#define OR_TR_NTETRAD 62
#define OR_TR_NCOORD 78
#define OR_TR_ECON 94
#define OR_TR_ECOV 110
int or_trace_pixel(const or_params_t *p, const or_camera_t *cam, long i, long j, const double *rub,
                   int cap, double *out, int *nstep_out);

int or_render_adaptive(const or_params_t *p, or_camera_t *cam, int nx_min, int ny_min, double refine_abs,
                       double refine_rel, double refine_cut, int nearest, double *image, double *imageS,
                       double *taus, int *interp, double *Iavg, long long *traced, long long *interpolated,
                       int nthreads);

/* image (src/main.c) */
int or_setup_camera(const or_params_t *p, or_camera_t *cam);
int or_render(const or_params_t *p, or_camera_t *cam, double *image, double *imageS,
              double *taus, int *nsteps, int *flags, int nthreads);
int or_render_pixels(const or_params_t *p, const or_camera_t *cam, const int *pix,
                     int npix, double *out7, int *nsteps, int *flags, int nthreads);

#endif
