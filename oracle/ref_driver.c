/*
 * Driver for the reference build in oracle/_ref (test infrastructure).
 *
 * It is linked with the reference's own translation units, compiled
 * unmodified from /root/reference by oracle/Makefile (target `ref`):
 * src/{coordinates,geometry,geodesics,model_geodesics,tetrads,model_tetrads,
 * grid,simcoords,koral_coords,debug_tools,utils,hdf5_utils,ipolarray,
 * radiation,par,dict}.c and one model plugin (model/<name>/model.c).
 * src/model_radiation.c and src/symphony/ include GSL headers, which this
 * image lacks, so they are not built; nothing here calls into them
 * (jar_calc/get_jkinv are referenced only by integrate_emission/evolve_N,
 * which this driver never calls).
 *
 * Usage: ref_driver <cmd> <in.bin> <out.bin> [--key=value ...]
 * Keys are model keys (passed to the plugin's try_set_model_parameter) or
 * the camera/geodesic keys handled here.  Inputs and outputs are flat
 * float64 arrays; the first input double is the record count.
 */
#include "decs.h"
#include "coordinates.h"
#include "geometry.h"
#include "geodesics.h"
#include "grid.h"
#include "ipolarray.h"
#include "model.h"
#include "model_geodesics.h"
#include "model_tetrads.h"
#include "par.h"
#include "radiation.h"
#include "tetrads.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* globals the reference's main.c owns (src/main.c:62-64) */
double tf = 0.;
Params params = {0};

static double *read_all(const char *fn, size_t *n)
{
  FILE *f = fopen(fn, "rb");
  if (!f) { perror(fn); exit(2); }
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  double *d = malloc(sz);
  if (fread(d, 1, sz, f) != (size_t)sz) { perror("read"); exit(2); }
  fclose(f);
  *n = sz / sizeof(double);
  return d;
}

static FILE *fout;
static void put(const double *v, size_t n) { fwrite(v, sizeof(double), n, fout); }
static void put1(double v) { put(&v, 1); }

static void cnorm(double complex N[NDIM][NDIM], const double *in)
{
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) N[i][j] = in[2 * (4 * i + j)] + I * in[2 * (4 * i + j) + 1];
}
static void cput(double complex N[NDIM][NDIM])
{
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      put1(creal(N[i][j]));
      put1(cimag(N[i][j]));
    }
}

int main(int argc, char *argv[])
{
  if (argc < 4) {
    fprintf(stderr, "usage: %s cmd in.bin out.bin [--key=value...]\n", argv[0]);
    return 2;
  }
  const char *cmd = argv[1];
  /* camera defaults as src/par.c:43-101 */
  params.rcam = 1000.;
  params.thetacam = 90.;
  params.nx = 160;
  params.ny = 160;
  params.eps = 0.01;
  params.maxnstep = 50000;
  params.dsource = DM87_PC;
  params.target_nturns = -1;
  for (int i = 4; i < argc; i++) {
    if (strncmp(argv[i], "--", 2) != 0) continue;
    char word[256], value[256];
    char *eq = strchr(argv[i], '=');
    if (!eq) continue;
    size_t wl = eq - argv[i] - 2;
    memcpy(word, argv[i] + 2, wl);
    word[wl] = 0;
    strncpy(value, eq + 1, 255);
    value[255] = 0;
    set_by_word_val(word, value, "rcam", &params.rcam, TYPE_DBL);
    set_by_word_val(word, value, "thetacam", &params.thetacam, TYPE_DBL);
    set_by_word_val(word, value, "phicam", &params.phicam, TYPE_DBL);
    set_by_word_val(word, value, "rotcam", &params.rotcam, TYPE_DBL);
    set_by_word_val(word, value, "nx", &params.nx, TYPE_INT);
    set_by_word_val(word, value, "ny", &params.ny, TYPE_INT);
    set_by_word_val(word, value, "eps", &params.eps, TYPE_DBL);
    set_by_word_val(word, value, "maxnstep", &params.maxnstep, TYPE_INT);
    set_by_word_val(word, value, "freqcgs", &params.freqcgs, TYPE_DBL);
    set_by_word_val(word, value, "dsource", &params.dsource, TYPE_DBL);
    set_by_word_val(word, value, "fovx_dsource", &params.fovx_dsource, TYPE_DBL);
    set_by_word_val(word, value, "fovy_dsource", &params.fovy_dsource, TYPE_DBL);
    set_by_word_val(word, value, "dx", &params.dx, TYPE_DBL);
    set_by_word_val(word, value, "dy", &params.dy, TYPE_DBL);
    set_by_word_val(word, value, "old_centering", &params.old_centering, TYPE_INT);
    set_by_word_val(word, value, "xoff", &params.xoff, TYPE_DBL);
    set_by_word_val(word, value, "yoff", &params.yoff, TYPE_DBL);
    set_by_word_val(word, value, "subring_dtheta", &params.subring_dtheta, TYPE_INT);
    try_set_model_parameter(word, value);
  }
  double tA, tB;
  init_model(&tA, &tB);

  size_t nin;
  double *in = read_all(argv[2], &nin);
  fout = fopen(argv[3], "wb");
  int n = (int)in[0];
  const double *rec = in + 1;

  if (!strcmp(cmd, "geom")) {
    put1(metric); put1(a); put1(hslope); put1(poly_xt); put1(poly_alpha); put1(mks_smooth);
    put1(poly_norm);
    put(startx, 4); put(stopx, 4); put(dx, 4); put(cstartx, 4); put(cstopx, 4);
    put1(Rin); put1(Rout); put1(Rh); put1(rmin_geo); put1(rmax_geo);
    put1(N1); put1(N2); put1(N3); put1(L_unit);
  } else if (!strcmp(cmd, "metric")) {
    for (int q = 0; q < n; q++) {
      double X[4], g[4][4], gc[4][4], conn[4][4][4], r, th;
      memcpy(X, rec + 4 * q, sizeof(X));
      gcov_func(X, g);
      gcon_func(g, gc);
      get_connection(X, conn);
      bl_coord(X, &r, &th);
      put(&g[0][0], 16); put(&gc[0][0], 16); put(&conn[0][0][0], 64); put1(r); put1(th);
      put1(gdet_func(g));
    }
  } else if (!strcmp(cmd, "trace")) {
    /* record = X[4] K[4] Xcam[4]; output per ray: nstep, then nstep+1 rows of
     * nturns dl X Kcon Xhalf Kconhalf */
    struct of_traj *traj = calloc(params.maxnstep, sizeof(struct of_traj));
    for (int q = 0; q < n; q++) {
      double X[4], K[4], Xc[4];
      memcpy(X, rec + 12 * q, sizeof(X));
      memcpy(K, rec + 12 * q + 4, sizeof(K));
      memcpy(Xc, rec + 12 * q + 8, sizeof(Xc));
      memset(traj, 0, params.maxnstep * sizeof(struct of_traj));
      int ns = trace_geodesic(X, K, traj, params.eps, params.maxnstep, Xc, &params, 0);
      put1(ns);
      for (int s = 0; s <= ns; s++) {
        put1(traj[s].nturns); put1(traj[s].dl); put(traj[s].X, 4); put(traj[s].Kcon, 4);
        put(traj[s].Xhalf, 4); put(traj[s].Kconhalf, 4);
      }
    }
    free(traj);
  } else if (!strcmp(cmd, "fluid")) {
    for (int q = 0; q < n; q++) {
      double X[4], K[4], Uc[4], Ul[4], Bc[4], Bl[4];
      memcpy(X, rec + 8 * q, sizeof(X));
      memcpy(K, rec + 8 * q + 4, sizeof(K));
      get_model_fourv(X, K, Uc, Ul, Bc, Bl);
      put(Uc, 4); put(Ul, 4); put(Bc, 4); put(Bl, 4);
      put1(get_model_ne(X)); put1(get_model_thetae(X)); put1(get_model_b(X));
      put1(get_model_sigma(X)); put1(radiating_region(X));
    }
  } else if (!strcmp(cmd, "trace_fields")) {
    /* the per-point model values dump_var_along records (src/io.c:357-371):
     * record X[4] K[4]; out: prims[8] sigma beta nu b_k_angle */
    for (int q = 0; q < n; q++) {
      double X[4], K[4], Uc[4], Ul[4], Bc[4], Bl[4], p8[8] = {0};
      memcpy(X, rec + 8 * q, sizeof(X));
      memcpy(K, rec + 8 * q + 4, sizeof(K));
      get_model_primitives(X, p8);
      get_model_fourv(X, K, Uc, Ul, Bc, Bl);
      put(p8, 8); put1(get_model_sigma(X)); put1(get_model_beta(X));
      put1(get_fluid_nu(K, Ul)); put1(get_bk_angle(X, K, Ul, Bc, Bl));
    }
  } else if (!strcmp(cmd, "ptvec")) {
    /* parallel_transport_vector (src/ipolarray.c:297-320): record Xm Km Ni Nm
     * dl; out Nf */
    for (int q = 0; q < n; q++) {
      const double *r0 = rec + 17 * q;
      double Xm[4], Km[4], Ni[4], Nm[4], Nf[4], X0[4] = {0}, K0[4] = {0};
      memcpy(Xm, r0, sizeof(Xm));
      memcpy(Km, r0 + 4, sizeof(Km));
      memcpy(Ni, r0 + 8, sizeof(Ni));
      memcpy(Nm, r0 + 12, sizeof(Nm));
      parallel_transport_vector(X0, Xm, X0, K0, Km, K0, Ni, Nm, Nf, r0[16]);
      put(Nf, 4);
    }
  } else if (!strcmp(cmd, "tetrad")) {
    for (int q = 0; q < n; q++) {
      double U[4], K[4], B[4], X[4], g[4][4], Ec[4][4], El[4][4];
      memcpy(U, rec + 16 * q, sizeof(U));
      memcpy(K, rec + 16 * q + 4, sizeof(K));
      memcpy(B, rec + 16 * q + 8, sizeof(B));
      memcpy(X, rec + 16 * q + 12, sizeof(X));
      gcov_func(X, g);
      int fl = make_plasma_tetrad(U, K, B, g, Ec, El);
      put(&Ec[0][0], 16); put(&El[0][0], 16); put1(fl);
    }
  } else if (!strcmp(cmd, "camtetrad")) {
    for (int q = 0; q < n; q++) {
      double X[4], Ec[4][4], El[4][4];
      memcpy(X, rec + 4 * q, sizeof(X));
      int fl = make_camera_tetrad(X, Ec, El);
      put(&Ec[0][0], 16); put(&El[0][0], 16); put1(fl);
      fl = make_camera_tetrad_old(X, Ec, El);
      put(&Ec[0][0], 16); put(&El[0][0], 16); put1(fl);
    }
  } else if (!strcmp(cmd, "push_polar")) {
    const int R = 4 + 4 + 32 + 32 + 1;
    for (int q = 0; q < n; q++) {
      const double *r0 = rec + R * q;
      double Xm[4], Km[4], X0[4] = {0}, K0[4] = {0};
      double complex Ni[4][4], Nm[4][4], Nf[4][4];
      memcpy(Xm, r0, sizeof(Xm));
      memcpy(Km, r0 + 4, sizeof(Km));
      cnorm(Ni, r0 + 8);
      cnorm(Nm, r0 + 40);
      push_polar(X0, Xm, X0, K0, Km, K0, Ni, Nm, Nf, r0[72]);
      cput(Nf);
    }
  } else if (!strcmp(cmd, "approx")) {
    for (int q = 0; q < n; q++) {
      const double *r0 = rec + 6 * q;
      double tau = 0.;
      double If = approximate_solve(r0[0], r0[1], r0[2], r0[3], r0[4], r0[5], &tau);
      put1(If); put1(tau);
    }
  } else if (!strcmp(cmd, "project")) {
    const int R = 4 + 32 + 1;
    for (int q = 0; q < n; q++) {
      const double *r0 = rec + R * q;
      double X[4], Kd[4] = {0}, SI, SQ, SU, SV;
      double complex N[4][4];
      memcpy(X, r0, sizeof(X));
      cnorm(N, r0 + 4);
      project_N(X, Kd, N, &SI, &SQ, &SU, &SV, r0[36]);
      put1(SI); put1(SQ); put1(SU); put1(SV);
    }
  } else if (!strcmp(cmd, "initxk")) {
    /* camera set-up of src/main.c:126-166, then init_XK and K *= freq */
    double Xcam[4];
    double freq = params.freqcgs * HPL / (ME * CL * CL);
    params.rotcam *= M_PI / 180.;
    native_coord(params.rcam, params.thetacam, params.phicam, Xcam);
    double Dsource = params.dsource * PC;
    double fov_to_d = Dsource / L_unit / MUAS_PER_RAD;
    if (params.fovx_dsource != 0.0) {
    } else if (params.dx != 0.0) {
      params.fovx_dsource = params.dx / fov_to_d;
      params.fovy_dsource = params.dy / fov_to_d;
    } else {
      params.fovx_dsource = 160.;
      params.fovy_dsource = 160.;
    }
    double DX = params.fovx_dsource * fov_to_d, DY = params.fovy_dsource * fov_to_d;
    double fovx = DX / params.rcam, fovy = DY / params.rcam;
    for (int q = 0; q < n; q++) {
      double X[4], K[4];
      init_XK((long)rec[2 * q], (long)rec[2 * q + 1], params.nx, params.ny, Xcam, params, fovx,
              fovy, X, K);
      for (int mu = 0; mu < 4; mu++) K[mu] *= freq;
      put(Xcam, 4); put(X, 4); put(K, 4);
    }
  } else {
    fprintf(stderr, "unknown command %s\n", cmd);
    return 2;
  }
  fclose(fout);
  free(in);
  return 0;
}
