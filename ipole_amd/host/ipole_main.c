/*
 * ipole host over the ipgpu C-ABI: the reference's command line, par files,
 * iharm dump loading, camera set-up, image statistics and HDF5 image layout,
 * with the pixel loop on the GPU.
 *
 * Follows src/main.c (fast light, no adaptive refinement), src/par.c and the
 * iharm plugin's set-up (model/iharm/model.c).  The isothermal sphere of
 * model/sphere/model.c is selected with `--model=sphere`: the reference
 * chooses its one model at compile time (makefile MODEL=...), this host at
 * run time.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "h5io.h"
#include "ipgpu.h"

/* src/constants.h */
#define HPL (6.6260693e-27)
#define ME (9.1093826e-28)
#define MP (1.67262171e-24)
#define CL (2.99792458e10)
#define GNEWT (6.6742e-8)
#define MSUN (1.989e33)
#define PC (3.085678e18)
#define JY (1.e-23)
#define MUAS_PER_RAD (2.06265e11)
#define SIGMA_THOMSON (0.665245873e-24)
#define DM87_PC (16.9e6)
#define VERSION_STRING "ipole-beta-1.5"

typedef struct {
  /* src/par.h:14-73 */
  double rcam, thetacam, phicam, rotcam, dx, dy, fovx_dsource, fovy_dsource, dsource, freqcgs;
  int nx, ny, old_centering, add_ppm, qu_conv, quench_output, only_unpolarized, perform_check;
  int emission_type, stokes_floors, isolate_counterjet, target_nturns, subring_dtheta;
  double eps, xoff, yoff;
  int maxnstep, trace, histo, histo_polar, nx_min, ny_min;
  int trace_stride, trace_i, trace_j, nearest_neighbor;
  double refine_abs, refine_rel, refine_cut;
  char outf[2048], trace_outf[2048];
  /* model keys: model/iharm/model.c:130-167, model/sphere/model.c:33-42 */
  char model[32], dump[2048], kharma_format[8];
  double MBH, M_unit, trat_small, trat_large, beta_crit, sigma_cut, sigma_cut_high;
  double rmax_geo, rmin_geo, polar_cut, target_mdot;
  int reverse_field;
  double sph_R0, sph_tau0, sph_Thetae0, sph_beta0, sph_gam, sph_tp_over_te, sph_MBH;
  int have_MBH;
  int transfer_fp32; /* GPU library option, no reference counterpart */
} params_t;

static void defaults(params_t *p)
{
  memset(p, 0, sizeof(*p));
  /* src/par.c:47-100 */
  p->emission_type = 4;
  p->rcam = 1000.;
  p->thetacam = 90.;
  p->nx = 160;
  p->ny = 160;
  p->eps = 0.01;
  p->maxnstep = 50000;
  p->dsource = DM87_PC;
  p->target_nturns = -1;
  p->nx_min = -1;
  p->ny_min = -1;
  p->refine_rel = 1.e-1;
  p->refine_abs = 1.e-5;
  p->refine_cut = 0.0;
  strcpy(p->outf, "image.h5");
  p->trace_stride = 1;
  p->trace_i = -1;
  p->trace_j = -1;
  strcpy(p->trace_outf, "trace.h5");
  /* model defaults */
  strcpy(p->model, "iharm");
  strcpy(p->dump, "dump.h5");
  strcpy(p->kharma_format, "new");
  p->MBH = 6.2e9;
  p->trat_small = 1.;
  p->trat_large = 40.;
  p->beta_crit = 1.;
  p->sigma_cut = 1.;
  p->sigma_cut_high = -1.;
  p->rmax_geo = 100.;
  p->rmin_geo = 1.;
  p->polar_cut = -1;
  p->sph_R0 = 100.;
  p->sph_tau0 = 1.e-5;
  p->sph_Thetae0 = 10.;
  p->sph_beta0 = 20.;
  p->sph_gam = 13. / 9;
  p->sph_tp_over_te = 3.;
  p->sph_MBH = 4.1e6;
}

/* src/par.c:235-253 */
static void set_d(const char *w, const char *v, const char *key, double *dst)
{
  if (!strcmp(w, key)) sscanf(v, "%lf", dst);
}
static void set_i(const char *w, const char *v, const char *key, int *dst)
{
  if (!strcmp(w, key)) sscanf(v, "%d", dst);
}
static void set_s(const char *w, const char *v, const char *key, char *dst)
{
  if (!strcmp(w, key)) sscanf(v, "%s", dst);
}

/* src/par.c:139-210 plus the model and radiation keys */
static void try_set(const char *w, const char *v, params_t *p)
{
  set_d(w, v, "freqcgs", &p->freqcgs);
  set_i(w, v, "add_ppm", &p->add_ppm);
  set_i(w, v, "qu_conv", &p->qu_conv);
  set_i(w, v, "quench_output", &p->quench_output);
  set_i(w, v, "only_unpolarized", &p->only_unpolarized);
  set_i(w, v, "emission_type", &p->emission_type);
  set_i(w, v, "transfer_fp32", &p->transfer_fp32);
  set_i(w, v, "stokes_floors", &p->stokes_floors);
  set_i(w, v, "counterjet", &p->isolate_counterjet);
  set_i(w, v, "old_centering", &p->old_centering);
  set_d(w, v, "rcam", &p->rcam);
  set_d(w, v, "thetacam", &p->thetacam);
  set_d(w, v, "phicam", &p->phicam);
  set_d(w, v, "rotcam", &p->rotcam);
  set_d(w, v, "dsource", &p->dsource);
  set_d(w, v, "fovx", &p->fovx_dsource);
  set_d(w, v, "fovy", &p->fovy_dsource);
  set_d(w, v, "fovx_dsource", &p->fovx_dsource);
  set_d(w, v, "fovy_dsource", &p->fovy_dsource);
  set_d(w, v, "fov", &p->fovx_dsource);
  set_d(w, v, "fov", &p->fovy_dsource);
  set_d(w, v, "dx", &p->dx);
  set_d(w, v, "dy", &p->dy);
  set_i(w, v, "nx", &p->nx);
  set_i(w, v, "ny", &p->ny);
  set_i(w, v, "nx_min", &p->nx_min);
  set_i(w, v, "ny_min", &p->ny_min);
  set_d(w, v, "refine_abs", &p->refine_abs);
  set_d(w, v, "refine_rel", &p->refine_rel);
  set_d(w, v, "refine_cut", &p->refine_cut);
  set_i(w, v, "use_nearest_neighbor", &p->nearest_neighbor);
  set_i(w, v, "target_nturns", &p->target_nturns);
  set_i(w, v, "subring_dtheta", &p->subring_dtheta);
  set_d(w, v, "eps", &p->eps);
  set_i(w, v, "maxnstep", &p->maxnstep);
  set_d(w, v, "xoff", &p->xoff);
  set_d(w, v, "yoff", &p->yoff);
  set_s(w, v, "outfile", p->outf);
  set_i(w, v, "trace", &p->trace);
  set_i(w, v, "trace_stride", &p->trace_stride);
  set_i(w, v, "trace_i", &p->trace_i);
  set_i(w, v, "trace_j", &p->trace_j);
  set_s(w, v, "trace_outf", p->trace_outf);
  set_i(w, v, "histo", &p->histo);
  set_i(w, v, "histo_polar", &p->histo_polar);
  /* model keys */
  set_s(w, v, "model", p->model);
  if (!strcmp(w, "MBH")) {
    sscanf(v, "%lf", &p->MBH);
    p->sph_MBH = p->MBH;
    p->have_MBH = 1;
  }
  set_d(w, v, "M_unit", &p->M_unit);
  set_d(w, v, "mdot", &p->target_mdot);
  set_s(w, v, "dump", p->dump);
  if (!strcmp(w, "kharma_format")) snprintf(p->kharma_format, sizeof(p->kharma_format), "%s", v);
  set_d(w, v, "trat_small", &p->trat_small);
  set_d(w, v, "trat_large", &p->trat_large);
  set_d(w, v, "sigma_cut", &p->sigma_cut);
  set_d(w, v, "sigma_cut_high", &p->sigma_cut_high);
  set_d(w, v, "beta_crit", &p->beta_crit);
  set_d(w, v, "rmax_geo", &p->rmax_geo);
  set_d(w, v, "rmin_geo", &p->rmin_geo);
  set_i(w, v, "reverse_field", &p->reverse_field);
  set_d(w, v, "polar_cut_deg", &p->polar_cut);
  set_d(w, v, "R0", &p->sph_R0);
  set_d(w, v, "tau0", &p->sph_tau0);
  set_d(w, v, "Thetae0", &p->sph_Thetae0);
  set_d(w, v, "beta0", &p->sph_beta0);
  set_d(w, v, "gam", &p->sph_gam);
  set_d(w, v, "tp_over_te", &p->sph_tp_over_te);
}

/* src/par.c:213-232 */
static void load_par(const char *fname, params_t *p)
{
  char line[256] = {0}, word[256] = {0}, value[256] = {0};
  FILE *fp = fopen(fname, "r");
  if (fp == NULL) {
    fprintf(stderr, "! unable to load parameter file '%s'. (%d: %s)\n", fname, errno, strerror(errno));
    exit(errno);
  }
  while (fgets(line, 255, fp) != NULL) {
    if (line[0] == '#') continue;
    word[0] = value[0] = 0;
    if (sscanf(line, "%s %s", word, value) < 1) continue;
    try_set(word, value, p);
  }
  fclose(fp);
}

/* src/par.c:43-136 */
static void load_par_from_argv(int argc, char *argv[], params_t *p)
{
  for (int i = 0; i < argc; ++i) {
    if (strcmp(argv[i], "-quench") == 0) p->quench_output = 1;
    else if (strcmp(argv[i], "-unpol") == 0) p->only_unpolarized = 1;
    else if (strcmp(argv[i], "-check") == 0) p->perform_check = 1;
    if (strlen(argv[i]) > 2 && argv[i][0] == '-' && argv[i][1] == '-') {
      char buf[4096];
      strncpy(buf, argv[i] + 2, sizeof(buf) - 1);
      buf[sizeof(buf) - 1] = 0;
      char *save = NULL;
      char *word = strtok_r(buf, "=", &save);
      char *value = strtok_r(NULL, "=", &save);
      if (value != NULL) try_set(word, value, p);
      else if (i == argc - 1 || (argv[i + 1][0] == '-' && argv[i + 1][1] == '-')) try_set(word, "1", p);
      else try_set(word, argv[i + 1], p);
    }
    if (strcmp(argv[i], "-par") == 0) {
      if (i == argc - 1) {
        fprintf(stderr, "! parameter filename must follow -par argument\n");
        exit(1);
      }
      load_par(argv[i + 1], p);
    }
  }
  if (p->histo != 0 || p->histo_polar != 0) {
    fprintf(stderr, "histogramming emission is nuanced. please either use the trace feature or "
                    "uncomment this exit() at your own risk.\n");
    exit(42);
  }
}

static double now(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

#define CHECK(call)                                                                \
  do {                                                                             \
    int rc_ = (call);                                                              \
    if (rc_ != IPGPU_OK) {                                                         \
      fprintf(stderr, "ipgpu: %s (%d)\n", ipgpu_last_error(ctx), rc_);             \
      exit(rc_ == IPGPU_ERR_MAXNSTEP ? -10 : -1);                                  \
    }                                                                              \
  } while (0)

int main(int argc, char *argv[])
{
  fprintf(stderr, "%s. githash: %s\n", VERSION_STRING, "ipole_amd");
  fprintf(stderr, "notes: %s\n\n", "MI355X pixel pipeline");
  double t_start = now();
  params_t par;
  defaults(&par);
  load_par_from_argv(argc, argv, &par);
  /* src/main.c:98-106 */
  if (par.nx_min < 0) {
    par.nx_min = par.nx;
    par.ny_min = par.ny;
  }
  int refine_level = 1;
  if (par.nx_min != par.nx || par.ny_min != par.ny) {
    if (par.nx_min < 2 || par.nx_min > par.nx || par.ny_min < 2 || par.ny_min > par.ny) {
      fprintf(stderr, "! nx_min, ny_min must lie in 2 .. nx, ny\n");
      exit(1);
    }
    refine_level = (int)(log2((par.nx - par.nx % 2) / (par.nx_min - par.nx_min % 2)) + 1);
  }
  if (par.target_mdot > 0) {
    fprintf(stderr, "! mdot normalization is not supported by this build; set M_unit\n");
    exit(1);
  }

  ipgpu_ctx *ctx = NULL;
  if (ipgpu_create(&ctx, 0) != IPGPU_OK) {
    fprintf(stderr, "ipgpu: no GPU context\n");
    return 1;
  }
  ipgpu_geom geom;
  ipgpu_model model;
  memset(&geom, 0, sizeof(geom));
  memset(&model, 0, sizeof(model));
  double L_unit, T_unit = 0, Thetae_unit = 0, dump_t = 0;
  ipole_dump_header kh_h;          /* the KHARMA dump's header, for the image file */
  ipole_kharma_info kh_ki;
  memset(&kh_h, 0, sizeof(kh_h));
  memset(&kh_ki, 0, sizeof(kh_ki));
  int is_iharm = strcmp(par.model, "sphere") != 0;
  const int is_kharma = !strcmp(par.model, "kharma");
  if (is_iharm) {
    /* init_model / init_iharm_grid / load_iharm_data, model/iharm/model.c:295-389, :848-1074 */
    /* iharm dumps (model/iharm/model.c:848-1074) or, with model kharma,
     * Parthenon files (model/kharma/model.c:772-1010): the same fluid model
     * after loading */
    ipole_dump_header h;
    ipole_kharma_info ki;
    memset(&ki, 0, sizeof(ki));
    snprintf(ki.format, sizeof(ki.format), "%s", par.kharma_format);
    ki.game = 1.333333333333333;
    ki.gamp = 1.666666666666667;
    int hrc = is_kharma ? ipole_h5_read_kharma_header(par.dump, &h, &ki) : ipole_h5_read_dump_header(par.dump, &h);
    if (hrc == -2) {
      fprintf(stderr, "Unknown coordinate system in %s\n", par.dump);
      exit(-1);
    }
    if (hrc != 0) {
      fprintf(stderr, "! unable to open file %s. Exiting!\n", par.dump);
      exit(-1);
    }
    if (!strncmp(h.metric, "MKS", 19)) geom.metric = IPGPU_METRIC_MKS;
    else if (!strncmp(h.metric, "FMKS", 19) || !strncmp(h.metric, "MMKS", 19)) geom.metric = IPGPU_METRIC_FMKS;
    else {
      fprintf(stderr, "File is in unknown metric %s.  Cannot continue.\n", h.metric);
      exit(-1);
    }
    geom.N1 = h.N1;
    geom.N2 = h.N2;
    geom.N3 = h.N3;
    geom.a = h.a;
    geom.hslope = h.hslope;
    geom.Rin = h.Rin;
    geom.Rout = h.Rout;
    if (geom.metric == IPGPU_METRIC_FMKS) {
      geom.poly_xt = h.poly_xt;
      geom.poly_alpha = h.poly_alpha;
      geom.mks_smooth = h.mks_smooth;
      geom.poly_norm = 0.5 * M_PI * 1. / (1. + 1. / (h.poly_alpha + 1.) * 1. / pow(h.poly_xt, h.poly_alpha));
    }
    for (int mu = 1; mu < 4; mu++) {
      geom.startx[mu] = h.startx[mu];
      geom.dx[mu] = h.dx[mu];
    }
    geom.rmax_geo = fmin(par.rmax_geo, geom.Rout);
    geom.rmin_geo = fmax(par.rmin_geo, geom.Rin);
    geom.stopx[0] = 1.;
    geom.stopx[1] = geom.startx[1] + geom.N1 * geom.dx[1];
    geom.stopx[2] = geom.startx[2] + geom.N2 * geom.dx[2];
    geom.stopx[3] = geom.startx[3] + geom.N3 * geom.dx[3];
    geom.cstopx[1] = log(geom.Rout);
    geom.cstopx[2] = 1.0;
    geom.cstopx[3] = 2 * M_PI;
    geom.Rh = 1 + sqrt(1. - geom.a * geom.a);
    double MBH = par.MBH * MSUN;
    L_unit = GNEWT * MBH / (CL * CL);
    T_unit = L_unit / CL;
    model.kind = IPGPU_MODEL_IHARM;
    model.L_unit = L_unit;
    model.th_beg = (par.polar_cut >= 0) ? 0.0174 * par.polar_cut : 0.0174;
    model.sigma_cut = par.sigma_cut;
    model.sigma_cut_high = par.sigma_cut_high;
    fprintf(stderr, "MBH: %g [Msun]\n", MBH / MSUN);
    fprintf(stderr, "L,T,M units: %g [cm] %g [s] %g [g]\n", L_unit, T_unit, par.M_unit);
    size_t n = (size_t)h.N1 * h.N2 * h.N3 * h.n_prim;
    double *prims = malloc(n * sizeof(double));
    if (!prims || (is_kharma ? ipole_h5_read_kharma_prims(par.dump, &h, &ki, prims)
                             : ipole_h5_read_dump_prims(par.dump, &h, prims)) != 0) {
      fprintf(stderr, "! unable to read prims from %s. Exiting!\n", par.dump);
      exit(-1);
    }
    if (par.reverse_field) /* model/iharm/model.c:1929-1941 */
      for (size_t c = 0; c < (size_t)h.N1 * h.N2 * h.N3; c++)
        for (int l = 5; l < 8; l++) prims[c * h.n_prim + l] *= -1.0;
    dump_t = h.t;
    kh_h = h;
    kh_ki = ki;
    ipgpu_units u = {par.M_unit, par.MBH, par.trat_small, par.trat_large, par.beta_crit,
                     h.gam, ki.game, ki.gamp};
    CHECK(ipgpu_set_geometry(ctx, &geom));
    CHECK(ipgpu_set_model(ctx, &model));
    CHECK(ipgpu_upload_prims(ctx, prims, h.n_prim, &u));
    free(prims);
  } else {
    /* model/sphere/model.c:49-117 */
    double MBH = par.sph_MBH * MSUN;
    L_unit = GNEWT * MBH / (CL * CL);
    model.kind = IPGPU_MODEL_SPHERE;
    model.L_unit = L_unit;
    model.sph_R0 = par.sph_R0;
    model.sph_Ne0 = par.sph_tau0 / SIGMA_THOMSON / par.sph_R0 / L_unit;
    double gam = par.sph_gam;
    double THETAE_UNIT = MP / ME * (gam - 1.) / (1. + par.sph_tp_over_te);
    model.sph_B0 = CL * sqrt(8 * M_PI * (gam - 1.) * (MP + ME) / par.sph_beta0)
        * sqrt(model.sph_Ne0 * par.sph_Thetae0) / sqrt(THETAE_UNIT);
    model.sph_Thetae0 = par.sph_Thetae0;
    model.sph_sigma = par.sph_Thetae0 * par.sph_tp_over_te / par.sph_beta0;
    model.sph_beta = par.sph_beta0;
    geom.metric = IPGPU_METRIC_EMINKOWSKI;
    geom.N1 = 300;
    geom.N2 = 100;
    geom.N3 = 1;
    geom.dx[0] = 0.1;
    geom.dx[1] = par.sph_R0 / geom.N1;
    geom.dx[2] = M_PI / geom.N2;
    geom.dx[3] = 2. * M_PI / geom.N3;
    geom.stopx[0] = 1.;
    geom.stopx[1] = geom.startx[1] + geom.N1 * geom.dx[1];
    geom.stopx[2] = geom.startx[2] + geom.N2 * geom.dx[2];
    geom.stopx[3] = geom.startx[3] + geom.N3 * geom.dx[3];
    for (int mu = 0; mu < 4; mu++) {
      geom.cstartx[mu] = geom.startx[mu];
      geom.cstopx[mu] = geom.stopx[mu];
    }
    geom.rmin_geo = 0;
    geom.rmax_geo = fmax(100., par.sph_R0);
    fprintf(stderr, "Running with isothermal sphere model.\n");
    CHECK(ipgpu_set_geometry(ctx, &geom));
    CHECK(ipgpu_set_model(ctx, &model));
  }

  /* camera, src/main.c:125-174 */
  double freq = par.freqcgs * HPL / (ME * CL * CL);
  if (freq == 0) {
    fprintf(stderr, "Frequency cannot be zero. Quitting!\n");
    exit(1);
  }
  double rotcam = par.rotcam * M_PI / 180.;
  double Xcam[4];
  CHECK(ipgpu_native_coord(&geom, par.rcam, par.thetacam, par.phicam, Xcam));
  fprintf(stderr, "Xcam[] = %e %e %e %e\n", Xcam[0], Xcam[1], Xcam[2], Xcam[3]);
  double Dsource = par.dsource * PC;
  double fov_to_d = Dsource / L_unit / MUAS_PER_RAD;
  if (par.fovx_dsource != 0.0) {
  } else if (par.dx != 0.0) {
    par.fovx_dsource = par.dx / fov_to_d;
    par.fovy_dsource = par.dy / fov_to_d;
  } else {
    fprintf(stderr, "No FOV was specified. Using default 160muas!\n");
    par.fovx_dsource = 160.;
    par.fovy_dsource = 160.;
  }
  double DX = par.fovx_dsource * fov_to_d, DY = par.fovy_dsource * fov_to_d;
  par.dx = DX;
  par.dy = DY;
  double fovx = DX / par.rcam, fovy = DY / par.rcam;
  double scale = (DX * L_unit / par.nx) * (DY * L_unit / par.ny) / (Dsource * Dsource) / JY;
  fprintf(stderr, "intensity [cgs] to flux per pixel [Jy] conversion: %g\n", scale);
  fprintf(stderr, "Dsource: %g [cm]\n", Dsource);
  fprintf(stderr, "FOVx, FOVy: %g %g [GM/c^2]\n", DX, DY);
  if (refine_level > 1) {
    fprintf(stderr, "Resolution: %dx%d, refined up to %dx%d (%d levels)\n", par.nx_min, par.ny_min, par.nx,
            par.ny, refine_level);
    fprintf(stderr,
            "Refinement when relative error is >%g%% or absolute error is >%g%% of estimated total flux, for "
            "pixel brightness > %f%% of average\n",
            par.refine_rel * 100, par.refine_abs * 100, par.refine_cut * 100);
  } else {
    fprintf(stderr, "Resolution: %dx%d\n", par.nx, par.ny);
  }

  ipgpu_image im;
  memset(&im, 0, sizeof(im));
  im.nx = par.nx;
  im.ny = par.ny;
  memcpy(im.Xcam, Xcam, sizeof(Xcam));
  im.fovx = fovx;
  im.fovy = fovy;
  im.freq = freq;
  im.freqcgs = par.freqcgs;
  im.rotcam = rotcam;
  im.xoff = par.xoff;
  im.yoff = par.yoff;
  im.eps = par.eps;
  im.maxnstep = par.maxnstep;
  im.old_centering = par.old_centering;
  im.only_unpolarized = par.only_unpolarized;
  im.qu_conv = par.qu_conv;
  im.stokes_floors = par.stokes_floors;
  im.target_nturns = par.target_nturns;
  im.subring_dtheta = par.subring_dtheta;
  im.counterjet = par.isolate_counterjet;
  im.emission_type = par.emission_type;
  im.transfer_fp32 = par.transfer_fp32;

  size_t npix = (size_t)par.nx * par.ny;
  double *image = calloc(npix, sizeof(double)), *taus = calloc(npix, sizeof(double));
  double *imageS = par.only_unpolarized ? NULL : calloc(npix * 5, sizeof(double));
  int *nstep = calloc(npix, sizeof(int)), *flag = calloc(npix, sizeof(int));
  if (!image || !taus || !nstep || !flag || (!par.only_unpolarized && !imageS)) {
    fprintf(stderr, "Could not allocate image memory!");
    exit(-1);
  }
  double init_time = now() - t_start;
  int rc;
  if (refine_level > 1) {
    /* src/main.c:505-792 on the device: the coarse pass, then one pass per
     * level over the pixels the error estimate selects */
    ipgpu_refine rf = {par.nx_min, par.ny_min, par.refine_abs, par.refine_rel, par.refine_cut,
                       par.nearest_neighbor};
    rc = ipgpu_render_adaptive(ctx, &im, &rf, image, imageS, taus, nstep, flag, NULL);
    if (rc == IPGPU_OK) {
      int lv;
      double Iavg;
      long long tr[IPGPU_MAX_LEVELS], ip[IPGPU_MAX_LEVELS];
      ipgpu_adaptive_stats(ctx, &lv, &Iavg, tr, ip);
      fprintf(stderr, "\n\n");
      const int nxi = (par.nx % 2 == 0) ? par.nx + 1 : par.nx;
      const int nxmi = (par.nx % 2 == 0) ? par.nx_min + 1 : par.nx_min;
      const long sp = (nxi - 1) / (nxmi - 1);
      for (int l = 1; l < lv; l++) {
        long ns = sp >> l;
        fprintf(stderr, "Refining level %d of %d, spacing %ld,%ld\n", l + 1, lv, ns, ns);
        long nxl = par.nx / ns, nyl = par.ny / ns;
        fprintf(stderr, "\n%lld of %ld (%f%%) of computed pixels at %ldx%ld were interpolated\n\n", ip[l],
                nxl * nyl, ((double)ip[l]) / (nxl * nyl) * 100, nxl, nyl);
      }
    }
  } else {
    rc = ipgpu_render(ctx, &im, image, imageS, taus, nstep, flag);
  }
  if (rc == IPGPU_ERR_MAXNSTEP) {
    int bi, bj;
    ipgpu_bad_pixel(ctx, &bi, &bj);
    fprintf(stderr, "\nMaxNStep exceeded in pixel %d %d!\n", bi, bj);
    exit(-10);
  } else if (rc == IPGPU_ERR_NAN) {
    fprintf(stderr, "NaN in image! Exiting.\n");
    exit(-1);
  }
  CHECK(rc);
  /* src/main.c:855-861 */
  for (size_t q = 0; q < npix; q++) {
    double Is = par.only_unpolarized ? 0. : imageS[q * 5] / pow(par.freqcgs, 3);
    if (flag[q] != 0 || Is < -1.e-10) {
      fprintf(stderr, "\nOddity in pixel %zu %zu (flag %d):\n", q / par.ny, q % par.ny, flag[q]);
      fprintf(stderr, "nstep: %d\n", nstep[q]);
    }
  }

  /* print_image_stats, src/main.c:1001-1042 */
  {
    double Ftot = 0., Ftot_unpol = 0., Imax = 0., Iavg = 0., Qtot = 0., Utot = 0., Vtot = 0.;
    size_t imax = 0, jmax = 0, nx = par.nx, ny = par.ny;
    for (size_t i = 0; i < nx; i++)
      for (size_t j = 0; j < ny; j++) {
        Ftot_unpol += image[i * ny + j] * scale;
        if (!par.only_unpolarized) {
          Ftot += imageS[(i * ny + j) * 5 + 0] * scale;
          Iavg += imageS[(i * ny + j) * 5 + 0];
          Qtot += imageS[(i * ny + j) * 5 + 1] * scale;
          Utot += imageS[(i * ny + j) * 5 + 2] * scale;
          Vtot += imageS[(i * ny + j) * 5 + 3] * scale;
          if (imageS[(i * ny + j) * 5 + 0] > Imax) {
            imax = i;
            jmax = j;
            Imax = imageS[(i * ny + j) * 5 + 0];
          }
        }
      }
    fprintf(stderr, "\nscale = %e\n", scale);
    fprintf(stderr, "imax=%zu jmax=%zu Imax=%g Iavg=%g\n", imax, jmax, Imax, Iavg / (par.nx * par.ny));
    fprintf(stderr, "freq: %g Ftot: %g Jy (%g Jy unpol xfer) scale=%g\n", par.freqcgs, Ftot, Ftot_unpol, scale);
    fprintf(stderr, "nuLnu = %g erg/s\n", 4. * M_PI * Ftot * Dsource * Dsource * JY * par.freqcgs);
    double LPfrac = 100. * sqrt(Qtot * Qtot + Utot * Utot) / Ftot;
    double CPfrac = 100. * Vtot / Ftot;
    fprintf(stderr, "I,Q,U,V [Jy]: %g %g %g %g\n", Ftot, Qtot, Utot, Vtot);
    fprintf(stderr, "LP,CP [%%]: %g %g\n", LPfrac, CPfrac);
  }

  ipole_image_meta mt;
  {
    memset(&mt, 0, sizeof(mt));
    strcpy(mt.version, VERSION_STRING);
    strcpy(mt.githash, "ipole_amd");
    mt.nx = par.nx;
    mt.ny = par.ny;
    mt.only_unpolarized = par.only_unpolarized;
    mt.qu_conv = par.qu_conv;
    mt.model_kind = is_iharm ? 1 : 2;
    mt.reverse_field = par.reverse_field;
    mt.freqcgs = par.freqcgs;
    mt.scale = scale;
    mt.dsource = Dsource;
    mt.dx = par.dx;
    mt.dy = par.dy;
    mt.fovx_dsource = par.fovx_dsource;
    mt.fovy_dsource = par.fovy_dsource;
    mt.rcam = par.rcam;
    mt.thetacam = par.thetacam;
    mt.phicam = par.phicam;
    mt.rotcam = rotcam;
    mt.fovx = fovx;
    mt.fovy = fovy;
    mt.xoff = par.xoff;
    mt.yoff = par.yoff;
    memcpy(mt.Xcam, Xcam, sizeof(Xcam));
    mt.t = dump_t;
    mt.sigma_cut = par.sigma_cut;
    mt.trat_small = par.trat_small;
    mt.trat_large = par.trat_large;
    mt.beta_crit = par.beta_crit;
    mt.L_unit = L_unit;
    mt.M_unit = par.M_unit;
    mt.T_unit = T_unit;
    mt.Thetae_unit = Thetae_unit;
  }

  /* trace output of get_pixel, src/main.c:865-880, written with
   * dump_var_along's layout (src/io.c:279-543): pixel (trace_i, trace_j)
   * alone, or every trace_stride-th pixel in both directions */
  if (par.trace) {
    const int stride = par.trace_stride > 0 ? par.trace_stride : 1;
    const int single = !(par.trace_i < 0 || par.trace_j < 0);
    const int tnx = single ? 1 : par.nx / stride, tny = single ? 1 : par.ny / stride;
    double *rec = malloc(sizeof(double) * (size_t)par.maxnstep * IPGPU_TRACE_NF);
    int64_t tf = -1;
    if (!rec) {
      fprintf(stderr, "Could not allocate trace memory!");
      exit(-1);
    }
    for (int i = 0; i < par.nx; i++)
      for (int j = 0; j < par.ny; j++) {
        int ti, tj;
        if (single) {
          if (i != par.trace_i || j != par.trace_j) continue;
          ti = tj = 0;
        } else {
          if (i % stride != 0 || j % stride != 0) continue;
          ti = i / stride;
          tj = j / stride;
          if (ti >= tnx || tj >= tny) continue; /* outside the file's nx/stride x ny/stride */
        }
        if (tf < 0) {
          int existed = access(par.trace_outf, F_OK) != -1;
          tf = ipole_h5_trace_open(par.trace_outf, &mt, tnx, tny, par.maxnstep);
          if (tf < 0) {
            fprintf(stderr, "! could not open %s\n", par.trace_outf);
            exit(-1);
          }
          if (!existed && is_iharm) {
            ipole_h5_trace_close(tf);
            if (is_kharma) ipole_h5_write_kharma_fluid_header(par.trace_outf, &kh_h, &kh_ki);
            else ipole_h5_copy_fluid_header(par.dump, par.trace_outf);
            tf = ipole_h5_trace_open(par.trace_outf, &mt, tnx, tny, par.maxnstep);
          }
        }
        int ns = 0;
        CHECK(ipgpu_trace_pixel(ctx, &im, i, j, par.maxnstep, &ns, rec));
        if (ipole_h5_trace_write(tf, ti, tj, ns, rec) != 0) {
          fprintf(stderr, "! could not write the trace of pixel %d %d to %s\n", i, j, par.trace_outf);
          exit(-1);
        }
      }
    if (tf >= 0) ipole_h5_trace_close(tf);
    free(rec);
  }

  if (!par.quench_output) {
    if (ipole_h5_write_image(par.outf, &mt, image, imageS, taus) != 0) {
      fprintf(stderr, "! could not write %s\n", par.outf);
      exit(-1);
    }
    if (is_iharm && !is_kharma) ipole_h5_copy_fluid_header(par.dump, par.outf);
    if (is_kharma) ipole_h5_write_kharma_fluid_header(par.outf, &kh_h, &kh_ki);
  }
  ipgpu_destroy(ctx);
  double t = now() - t_start;
  fprintf(stderr, "Total wallclock time: %g s (%g s)\n\n", t, init_time);
  return 0;
}
