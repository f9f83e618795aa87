/*
 * HDF5 I/O of the host: iharm-format dumps (read and write) and the image
 * file layout of the reference (src/io.c:140-273, model/iharm/model.c:
 * 1254-1301).  Plain C over libhdf5; no GPU code.
 */
#include "h5io.h"

#include <ctype.h>
#include <hdf5.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "ipgpu.h"

static int w_scalar(hid_t loc, const char *name, hid_t type, const void *v)
{
  hid_t sp = H5Screate(H5S_SCALAR);
  hid_t ds = H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) {
    H5Sclose(sp);
    return -1;
  }
  herr_t e = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
  H5Dclose(ds);
  H5Sclose(sp);
  return e < 0 ? -1 : 0;
}

static int w_str(hid_t loc, const char *name, const char *s, size_t len)
{
  hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, len);
  char buf[256] = {0};
  strncpy(buf, s, sizeof(buf) - 1);
  int rc = w_scalar(loc, name, t, buf);
  H5Tclose(t);
  return rc;
}

static int w_d(hid_t loc, const char *name, double v) { return w_scalar(loc, name, H5T_IEEE_F64LE, &v); }
static int w_i(hid_t loc, const char *name, int v) { return w_scalar(loc, name, H5T_STD_I32LE, &v); }

static int w_array(hid_t loc, const char *name, int rank, const hsize_t *dims, hid_t type,
                   const void *data)
{
  hid_t sp = H5Screate_simple(rank, dims, NULL);
  hid_t ds = H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) {
    H5Sclose(sp);
    return -1;
  }
  herr_t e = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
  H5Dclose(ds);
  H5Sclose(sp);
  return e < 0 ? -1 : 0;
}

static int r_scalar(hid_t loc, const char *name, hid_t type, void *v)
{
  if (H5Lexists(loc, name, H5P_DEFAULT) <= 0) return -1;
  hid_t ds = H5Dopen2(loc, name, H5P_DEFAULT);
  if (ds < 0) return -1;
  herr_t e = H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
  H5Dclose(ds);
  return e < 0 ? -1 : 0;
}

/* Dump keys as model/iharm/model.c:860-1051 and :1895-1942 read them. */
int ipole_h5_write_dump(const char *path, const ipole_dump_header *h, const double *prims)
{
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = 0;
  hid_t hd = H5Gcreate2(f, "header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  rc |= w_str(hd, "version", "ipole_amd-synthetic", 20);
  rc |= w_str(hd, "metric", h->metric, 20);
  rc |= w_i(hd, "n1", h->N1);
  rc |= w_i(hd, "n2", h->N2);
  rc |= w_i(hd, "n3", h->N3);
  rc |= w_i(hd, "n_prim", h->n_prim);
  rc |= w_d(hd, "gam", h->gam);
  rc |= w_i(hd, "has_electrons", 0);
  hid_t geom = H5Gcreate2(hd, "geom", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  rc |= w_d(geom, "startx1", h->startx[1]);
  rc |= w_d(geom, "startx2", h->startx[2]);
  rc |= w_d(geom, "startx3", h->startx[3]);
  rc |= w_d(geom, "dx1", h->dx[1]);
  rc |= w_d(geom, "dx2", h->dx[2]);
  rc |= w_d(geom, "dx3", h->dx[3]);
  int fmks = (strcmp(h->metric, "FMKS") == 0 || strcmp(h->metric, "MMKS") == 0);
  hid_t mk = H5Gcreate2(geom, fmks ? "mmks" : "mks", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  rc |= w_d(mk, "a", h->a);
  rc |= w_d(mk, "hslope", h->hslope);
  rc |= w_d(mk, "Rin", h->Rin);
  rc |= w_d(mk, "Rout", h->Rout);
  if (fmks) {
    rc |= w_d(mk, "poly_xt", h->poly_xt);
    rc |= w_d(mk, "poly_alpha", h->poly_alpha);
    rc |= w_d(mk, "mks_smooth", h->mks_smooth);
  }
  H5Gclose(mk);
  H5Gclose(geom);
  H5Gclose(hd);
  rc |= w_d(f, "dump_cadence", h->dump_cadence);
  rc |= w_d(f, "t", h->t);
  hsize_t dims[4] = {(hsize_t)h->N1, (hsize_t)h->N2, (hsize_t)h->N3, (hsize_t)h->n_prim};
  rc |= w_array(f, "prims", 4, dims, H5T_IEEE_F64LE, prims);
  H5Fclose(f);
  return rc ? -1 : 0;
}

int ipole_h5_read_dump_header(const char *path, ipole_dump_header *h)
{
  memset(h, 0, sizeof(*h));
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = 0;
  hid_t s20 = H5Tcopy(H5T_C_S1);
  H5Tset_size(s20, 20);
  char metric[21] = {0};
  rc |= r_scalar(f, "/header/metric", s20, metric);
  H5Tclose(s20);
  strncpy(h->metric, metric, sizeof(h->metric) - 1);
  rc |= r_scalar(f, "/header/n1", H5T_NATIVE_INT, &h->N1);
  rc |= r_scalar(f, "/header/n2", H5T_NATIVE_INT, &h->N2);
  rc |= r_scalar(f, "/header/n3", H5T_NATIVE_INT, &h->N3);
  rc |= r_scalar(f, "/header/n_prim", H5T_NATIVE_INT, &h->n_prim);
  rc |= r_scalar(f, "/header/gam", H5T_NATIVE_DOUBLE, &h->gam);
  rc |= r_scalar(f, "/header/geom/startx1", H5T_NATIVE_DOUBLE, &h->startx[1]);
  rc |= r_scalar(f, "/header/geom/startx2", H5T_NATIVE_DOUBLE, &h->startx[2]);
  rc |= r_scalar(f, "/header/geom/startx3", H5T_NATIVE_DOUBLE, &h->startx[3]);
  rc |= r_scalar(f, "/header/geom/dx1", H5T_NATIVE_DOUBLE, &h->dx[1]);
  rc |= r_scalar(f, "/header/geom/dx2", H5T_NATIVE_DOUBLE, &h->dx[2]);
  rc |= r_scalar(f, "/header/geom/dx3", H5T_NATIVE_DOUBLE, &h->dx[3]);
  int fmks = (strncmp(h->metric, "FMKS", 19) == 0 || strncmp(h->metric, "MMKS", 19) == 0);
  const char *g = fmks ? "/header/geom/mmks/" : "/header/geom/mks/";
  char key[128];
#define RD(name, dst)                                                     \
  snprintf(key, sizeof(key), "%s%s", g, name);                            \
  rc |= r_scalar(f, key, H5T_NATIVE_DOUBLE, dst);
  RD("a", &h->a);
  RD("hslope", &h->hslope);
  snprintf(key, sizeof(key), "%sRin", g);
  if (H5Lexists(f, key, H5P_DEFAULT) > 0) {
    RD("Rin", &h->Rin);
    RD("Rout", &h->Rout);
  } else {
    RD("r_in", &h->Rin);
    RD("r_out", &h->Rout);
  }
  if (fmks) {
    RD("poly_xt", &h->poly_xt);
    RD("poly_alpha", &h->poly_alpha);
    RD("mks_smooth", &h->mks_smooth);
  }
#undef RD
  r_scalar(f, "/dump_cadence", H5T_NATIVE_DOUBLE, &h->dump_cadence);
  r_scalar(f, "/t", H5T_NATIVE_DOUBLE, &h->t);
  H5Fclose(f);
  return rc ? -1 : 0;
}

int ipole_h5_read_dump_prims(const char *path, const ipole_dump_header *h, double *prims)
{
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  hid_t ds = H5Dopen2(f, "/prims", H5P_DEFAULT);
  if (ds < 0) {
    H5Fclose(f);
    return -1;
  }
  herr_t e = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, prims);
  H5Dclose(ds);
  H5Fclose(f);
  (void)h;
  return e < 0 ? -1 : 0;
}

/* Image file, src/io.c:229-273 (dump) with the header of :140-191 and the
 * iharm model header of model/iharm/model.c:1254-1301. */
/* write_header, src/io.c:140-191, with the model's output_hdf5
 * (model/iharm/model.c:1254-1301): the /githash attribute, /header and, for
 * iharm, Mdot, MdotEdd and Ladv at the root. */
static int write_header(hid_t f, const ipole_image_meta *mt)
{
  int rc = 0;
  /* githash attribute on / */
  {
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, 20);
    char buf[20] = {0};
    strncpy(buf, mt->githash, sizeof(buf) - 1);
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t at = H5Acreate2(f, "githash", t, sp, H5P_DEFAULT, H5P_DEFAULT);
    if (at >= 0) {
      H5Awrite(at, t, buf);
      H5Aclose(at);
    }
    H5Sclose(sp);
    H5Tclose(t);
  }
  hid_t hd = H5Gcreate2(f, "header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  rc |= w_str(hd, "version", mt->version, 20);
  rc |= w_str(hd, "githash", mt->githash, 20);
  rc |= w_d(hd, "freqcgs", mt->freqcgs);
  rc |= w_d(hd, "scale", mt->scale);
  rc |= w_d(hd, "dsource", mt->dsource);
  rc |= w_str(hd, "evpa_0", mt->qu_conv == 0 ? "N" : "W", 2);
  hid_t cam = H5Gcreate2(hd, "camera", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  rc |= w_i(cam, "nx", mt->nx);
  rc |= w_i(cam, "ny", mt->ny);
  rc |= w_d(cam, "dx", mt->dx);
  rc |= w_d(cam, "dy", mt->dy);
  rc |= w_d(cam, "fovx_dsource", mt->fovx_dsource);
  rc |= w_d(cam, "fovy_dsource", mt->fovy_dsource);
  rc |= w_d(cam, "rcam", mt->rcam);
  rc |= w_d(cam, "thetacam", mt->thetacam);
  rc |= w_d(cam, "phicam", mt->phicam);
  rc |= w_d(cam, "rotcam", mt->rotcam);
  rc |= w_d(cam, "fovx", mt->fovx);
  rc |= w_d(cam, "fovy", mt->fovy);
  rc |= w_d(cam, "xoff", mt->xoff);
  rc |= w_d(cam, "yoff", mt->yoff);
  hsize_t d4 = 4;
  rc |= w_array(cam, "x", 1, &d4, H5T_IEEE_F64LE, mt->Xcam);
  H5Gclose(cam);
  if (mt->model_kind == 1) {
    /* model/iharm/model.c:1254-1301 (the /fluid_header copy is added by the
     * host when it has the dump file) */
    rc |= w_d(hd, "t", mt->t);
    rc |= w_d(hd, "sigma_cut", mt->sigma_cut);
    hid_t el = H5Gcreate2(hd, "electrons", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    rc |= w_d(el, "rlow", mt->trat_small);
    rc |= w_d(el, "rhigh", mt->trat_large);
    rc |= w_d(el, "beta_crit", mt->beta_crit);
    rc |= w_i(el, "type", 2);
    H5Gclose(el);
    rc |= w_i(hd, "field_config", mt->reverse_field);
    hid_t un = H5Gcreate2(hd, "units", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    rc |= w_d(un, "L_unit", mt->L_unit);
    rc |= w_d(un, "M_unit", mt->M_unit);
    rc |= w_d(un, "T_unit", mt->T_unit);
    rc |= w_d(un, "Thetae_unit", mt->Thetae_unit);
    H5Gclose(un);
    rc |= w_d(f, "Mdot", mt->Mdot);
    rc |= w_d(f, "MdotEdd", mt->MdotEdd);
    rc |= w_d(f, "Ladv", mt->Ladv);
  } else {
    hid_t un = H5Gcreate2(hd, "units", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    rc |= w_d(un, "L_unit", mt->L_unit);
    H5Gclose(un);
  }
  H5Gclose(hd);
  return rc;
}

int ipole_h5_write_image(const char *path, const ipole_image_meta *mt, const double *image,
                         const double *imageS, const double *taus)
{
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = write_header(f, mt);
  double Ftot_unpol = 0., Ftot = 0.;
  size_t nx = mt->nx, ny = mt->ny;
  for (size_t i = 0; i < nx; ++i)
    for (size_t j = 0; j < ny; ++j) {
      Ftot_unpol += image[i * ny + j] * mt->scale;
      if (!mt->only_unpolarized) Ftot += imageS[(i * ny + j) * 5 + 0] * mt->scale;
    }
  rc |= w_d(f, "Ftot_unpol", Ftot_unpol);
  if (!mt->only_unpolarized) rc |= w_d(f, "Ftot", Ftot);
  double nuLnu = 4. * 3.14159265358979323846 * Ftot * mt->dsource * mt->dsource * 1.e-23 * mt->freqcgs;
  rc |= w_d(f, "nuLnu", nuLnu);
  nuLnu = 4. * 3.14159265358979323846 * Ftot_unpol * mt->dsource * mt->dsource * 1.e-23 * mt->freqcgs;
  rc |= w_d(f, "nuLnu_unpol", nuLnu);
  hsize_t d2[2] = {nx, ny}, d3[3] = {nx, ny, 5};
  rc |= w_array(f, "unpol", 2, d2, H5T_IEEE_F64LE, image);
  rc |= w_array(f, "tau", 2, d2, H5T_IEEE_F64LE, taus);
  if (!mt->only_unpolarized) rc |= w_array(f, "pol", 3, d3, H5T_IEEE_F64LE, imageS);
  H5Fclose(f);
  return rc ? -1 : 0;
}

/* Copy /header of the dump into /fluid_header of the image file
 * (model/iharm/model.c:1259-1260, via an object copy). */
int ipole_h5_copy_fluid_header(const char *dump_path, const char *image_path)
{
  hid_t src = H5Fopen(dump_path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (src < 0) return -1;
  hid_t dst = H5Fopen(image_path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (dst < 0) {
    H5Fclose(src);
    return -1;
  }
  herr_t e = H5Ocopy(src, "/header", dst, "/fluid_header", H5P_DEFAULT, H5P_DEFAULT);
  H5Fclose(dst);
  H5Fclose(src);
  return e < 0 ? -1 : 0;
}

int ipole_h5_read_image(const char *path, int nx, int ny, double *image, double *imageS,
                        double *taus, double *Ftot, double *Ftot_unpol)
{
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = 0;
  (void)nx;
  (void)ny;
  if (image) rc |= r_scalar(f, "/unpol", H5T_NATIVE_DOUBLE, image);
  if (taus) rc |= r_scalar(f, "/tau", H5T_NATIVE_DOUBLE, taus);
  if (imageS) rc |= r_scalar(f, "/pol", H5T_NATIVE_DOUBLE, imageS);
  if (Ftot) r_scalar(f, "/Ftot", H5T_NATIVE_DOUBLE, Ftot);
  if (Ftot_unpol) rc |= r_scalar(f, "/Ftot_unpol", H5T_NATIVE_DOUBLE, Ftot_unpol);
  H5Fclose(f);
  return rc ? -1 : 0;
}

/* ------------------------------------------------------------------ traces */

/* Datasets of the trace file (src/io.c:445-520): name, offset of the field in
 * a record of IPGPU_TRACE_NF doubles, and the per-step shape (scalar, 4, 8 or
 * 4 x 4). */
static const struct {
  const char *name;
  int off, rank, d0, d1;
} TRACE_SETS[] = {
    {"b", IPGPU_TR_B, 0, 0, 0},
    {"ne", IPGPU_TR_NE, 0, 0, 0},
    {"thetae", IPGPU_TR_THETAE, 0, 0, 0},
    {"nu", IPGPU_TR_NU, 0, 0, 0},
    {"b_k_angle", IPGPU_TR_BK_ANGLE, 0, 0, 0},
    {"sigma", IPGPU_TR_SIGMA, 0, 0, 0},
    {"beta", IPGPU_TR_BETA, 0, 0, 0},
    {"dl", IPGPU_TR_DL, 0, 0, 0},
    {"r", IPGPU_TR_R, 0, 0, 0},
    {"th", IPGPU_TR_TH, 0, 0, 0},
    {"phi", IPGPU_TR_PHI, 0, 0, 0},
    {"j_unpol", IPGPU_TR_J_UNPOL, 0, 0, 0},
    {"k_unpol", IPGPU_TR_K_UNPOL, 0, 0, 0},
    {"I_unpol", IPGPU_TR_I_UNPOL, 0, 0, 0},
    {"X", IPGPU_TR_X, 1, 4, 0},
    {"Kcon", IPGPU_TR_KCON, 1, 4, 0},
    {"j_inv", IPGPU_TR_J_INV, 1, 4, 0},
    {"alpha_inv", IPGPU_TR_ALPHA_INV, 1, 4, 0},
    {"rho_inv", IPGPU_TR_RHO_INV, 1, 4, 0},
    {"stokes", IPGPU_TR_STOKES, 1, 4, 0},
    {"e2Constructed", IPGPU_TR_E2C, 1, 4, 0},
    {"e2ConstructedCov", IPGPU_TR_E2C_COV, 1, 4, 0},
    {"e2Transported", IPGPU_TR_E2T, 1, 4, 0},
    {"e2TransportedCov", IPGPU_TR_E2T_COV, 1, 4, 0},
    {"prims", IPGPU_TR_PRIMS, 1, 8, 0},
    {"ntetrad", IPGPU_TR_NTETRAD, 2, 4, 4},
    {"ncoord", IPGPU_TR_NCOORD, 2, 4, 4},
    {"Econ", IPGPU_TR_ECON, 2, 4, 4},
    {"Ecov", IPGPU_TR_ECOV, 2, 4, 4},
};
#define NTRACE_SETS ((int)(sizeof(TRACE_SETS) / sizeof(TRACE_SETS[0])))

int64_t ipole_h5_trace_open(const char *path, const ipole_image_meta *mt, int nx, int ny, int maxnstep)
{
  if (access(path, F_OK) != -1) { /* src/io.c:282-287: append to an existing file */
    hid_t f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    return f < 0 ? -1 : (int64_t)f;
  }
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = write_header(f, mt);
  /* nstep per pixel, then the per-step sets chunked by 200 steps; chunks no
   * pixel writes stay unallocated and read as 0 */
  hsize_t d2[2] = {(hsize_t)nx, (hsize_t)ny};
  hid_t sp = H5Screate_simple(2, d2, NULL);
  hid_t ds = H5Dcreate2(f, "nstep", H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) rc = -1;
  else H5Dclose(ds);
  H5Sclose(sp);
  const hsize_t ck = maxnstep < 200 ? (hsize_t)maxnstep : 200;
  for (int q = 0; q < NTRACE_SETS && rc == 0; q++) {
    hsize_t dims[5] = {(hsize_t)nx, (hsize_t)ny, (hsize_t)maxnstep, (hsize_t)TRACE_SETS[q].d0,
                       (hsize_t)TRACE_SETS[q].d1};
    hsize_t chunk[5] = {1, 1, ck, (hsize_t)TRACE_SETS[q].d0, (hsize_t)TRACE_SETS[q].d1};
    int rank = 3 + TRACE_SETS[q].rank;
    hid_t s3 = H5Screate_simple(rank, dims, NULL);
    hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
    H5Pset_chunk(pl, rank, chunk);
    hid_t d = H5Dcreate2(f, TRACE_SETS[q].name, H5T_IEEE_F64LE, s3, H5P_DEFAULT, pl, H5P_DEFAULT);
    if (d < 0) rc = -1;
    else H5Dclose(d);
    H5Pclose(pl);
    H5Sclose(s3);
  }
  if (rc) {
    H5Fclose(f);
    return -1;
  }
  return (int64_t)f;
}

int ipole_h5_trace_write(int64_t fid, int i, int j, int nstep, const double *rec)
{
  hid_t f = (hid_t)fid;
  int rc = 0;
  const size_t nsteps = (size_t)nstep + 1;
  {
    hid_t ds = H5Dopen2(f, "nstep", H5P_DEFAULT);
    if (ds < 0) return -1;
    hid_t fs = H5Dget_space(ds);
    hsize_t st[2] = {(hsize_t)i, (hsize_t)j}, cnt[2] = {1, 1};
    H5Sselect_hyperslab(fs, H5S_SELECT_SET, st, NULL, cnt, NULL);
    hid_t ms = H5Screate_simple(2, cnt, NULL);
    if (H5Dwrite(ds, H5T_NATIVE_INT, ms, fs, H5P_DEFAULT, &nstep) < 0) rc = -1;
    H5Sclose(ms);
    H5Sclose(fs);
    H5Dclose(ds);
  }
  double *buf = malloc(sizeof(double) * nsteps * 16);
  if (!buf) return -1;
  for (int q = 0; q < NTRACE_SETS && rc == 0; q++) {
    const int w = TRACE_SETS[q].rank == 0 ? 1 : (TRACE_SETS[q].rank == 1 ? TRACE_SETS[q].d0 : 16);
    for (size_t s = 0; s < nsteps; s++)
      for (int k = 0; k < w; k++) buf[s * w + k] = rec[s * IPGPU_TRACE_NF + TRACE_SETS[q].off + k];
    hid_t ds = H5Dopen2(f, TRACE_SETS[q].name, H5P_DEFAULT);
    if (ds < 0) {
      rc = -1;
      break;
    }
    hid_t fs = H5Dget_space(ds);
    const int rank = H5Sget_simple_extent_ndims(fs);
    hsize_t dims[5], st[5] = {(hsize_t)i, (hsize_t)j, 0, 0, 0}, cnt[5];
    H5Sget_simple_extent_dims(fs, dims, NULL);
    if ((hsize_t)i >= dims[0] || (hsize_t)j >= dims[1] || nsteps > dims[2]) rc = -1;
    cnt[0] = cnt[1] = 1;
    cnt[2] = nsteps;
    for (int r = 3; r < rank; r++) cnt[r] = dims[r];
    if (rc == 0) {
      H5Sselect_hyperslab(fs, H5S_SELECT_SET, st, NULL, cnt, NULL);
      hid_t ms = H5Screate_simple(rank, cnt, NULL);
      if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, buf) < 0) rc = -1;
      H5Sclose(ms);
    }
    H5Sclose(fs);
    H5Dclose(ds);
  }
  free(buf);
  return rc;
}

int ipole_h5_trace_close(int64_t fid) { return H5Fclose((hid_t)fid) < 0 ? -1 : 0; }

int ipole_h5_dataset_shape(const char *path, const char *name, int64_t *dims)
{
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int rank = -1;
  if (H5Lexists(f, name, H5P_DEFAULT) > 0) {
    hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    if (ds >= 0) {
      hid_t fs = H5Dget_space(ds);
      hsize_t d[16];
      rank = H5Sget_simple_extent_ndims(fs);
      if (rank > 16) rank = -1;
      else if (rank >= 0) {
        H5Sget_simple_extent_dims(fs, d, NULL);
        for (int r = 0; r < rank; r++) dims[r] = (int64_t)d[r];
      }
      H5Sclose(fs);
      H5Dclose(ds);
    }
  }
  H5Fclose(f);
  return rank;
}

int ipole_h5_read_doubles(const char *path, const char *name, double *out)
{
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int rc = r_scalar(f, name, H5T_NATIVE_DOUBLE, out);
  H5Fclose(f);
  return rc;
}

/* ----------------------------------------------------------------- KHARMA */

/* Attribute of a group as a double / int array (n values). */
static int r_attr(hid_t f, const char *group, const char *name, hid_t type, void *v)
{
  hid_t g = H5Gopen2(f, group, H5P_DEFAULT);
  if (g < 0) return -1;
  int rc = -1;
  if (H5Aexists(g, name) > 0) {
    hid_t a = H5Aopen(g, name, H5P_DEFAULT);
    if (a >= 0) {
      rc = H5Aread(a, type, v) < 0 ? -1 : 0;
      H5Aclose(a);
    }
  }
  H5Gclose(g);
  return rc;
}

/* The parameter file KHARMA stores in /Input "File" (model/kharma/model.c:
 * 407-500), fixed- or variable-length; malloc'ed, or NULL. */
static char *kharma_parfile(hid_t f)
{
  hid_t g = H5Gopen2(f, "Input", H5P_DEFAULT);
  if (g < 0) return NULL;
  char *out = NULL;
  hid_t a = H5Aopen(g, "File", H5P_DEFAULT);
  if (a >= 0) {
    hid_t t = H5Aget_type(a);
    if (H5Tis_variable_str(t) > 0) {
      char *vl = NULL;
      if (H5Aread(a, t, &vl) >= 0 && vl) {
        out = strdup(vl);
        H5free_memory(vl);
      }
    } else {
      size_t n = H5Tget_size(t);
      out = calloc(n + 1, 1);
      if (out && H5Aread(a, t, out) < 0) {
        free(out);
        out = NULL;
      }
    }
    H5Tclose(t);
    H5Aclose(a);
  }
  H5Gclose(g);
  return out;
}

/* Value of key in <block> of a Parthenon parameter file, with the lookup
 * rules of get_parameter_value (model/kharma/model.c:525-610): the block
 * runs to the next '<'; the key must stand alone (whitespace before,
 * whitespace or '=' after); the value follows '='.  kind: 'i', 'd' or 's'
 * (s: up to whitespace or '#', n bytes).  0 or -1 when absent. */
static int par_get(const char *text, const char *block, const char *key, char kind, void *v, size_t n)
{
  const char *lo = text, *hi = text + strlen(text);
  char tag[256];
  snprintf(tag, sizeof(tag), "<%s>", block);
  const char *b = strstr(text, tag);
  if (!b) return -1;
  lo = b;
  const char *nb = strchr(b + strlen(tag), '<');
  if (nb) hi = nb;
  const size_t kl = strlen(key);
  const char *p = lo, *at = NULL;
  while ((p = strstr(p, key)) != NULL && p < hi) {
    const int ok_before = (p == lo) || isspace((unsigned char)p[-1]);
    const char c = p[kl];
    const int ok_after = c == '\0' || isspace((unsigned char)c) || c == '=';
    if (ok_before && ok_after) {
      at = p;
      break;
    }
    p++;
  }
  if (!at) return -1;
  const char *eq = strchr(at, '=');
  if (!eq || eq >= hi) return -1;
  eq++;
  while (eq < hi && *eq && isspace((unsigned char)*eq)) eq++;
  if (kind == 'i') *(int *)v = atoi(eq);
  else if (kind == 'd') *(double *)v = atof(eq);
  else {
    char *s = (char *)v;
    size_t i = 0;
    while (eq + i < hi && eq[i] && !isspace((unsigned char)eq[i]) && eq[i] != '#' && i < n - 1) {
      s[i] = eq[i];
      i++;
    }
    s[i] = 0;
  }
  return 0;
}

/* read_parameters_and_allocate_memory (model/kharma/model.c:772-1010) for
 * the grid, coordinates and adiabatic indices, and the Info attributes of
 * load_kharma_data (:1267-1272, :1468-1470). */
int ipole_h5_read_kharma_header(const char *path, ipole_dump_header *h, ipole_kharma_info *ki)
{
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  memset(h, 0, sizeof(*h));
  char keep_format[8];
  memcpy(keep_format, ki->format, sizeof(keep_format));
  memset(ki, 0, sizeof(*ki));
  memcpy(ki->format, keep_format, sizeof(keep_format));
  if (!ki->format[0]) strcpy(ki->format, "new");
  int rc = 0;
  char *par = kharma_parfile(f);
  if (!par) rc = -1;
  rc |= r_attr(f, "Info", "NumMeshBlocks", H5T_NATIVE_INT, &ki->num_meshblocks);
  rc |= r_attr(f, "Info", "MeshBlockSize", H5T_NATIVE_INT, ki->mb);
  rc |= r_attr(f, "Info", "Time", H5T_NATIVE_DOUBLE, &h->t);
  H5Fclose(f);
  if (rc) {
    free(par);
    return -1;
  }
  char tr[32] = {0}, el[32] = {0};
  double x1max = 0., x2max = 0., x3max = 0.;
  rc |= par_get(par, "coordinates", "transform", 's', tr, sizeof(tr));
  if (!strncmp(tr, "fmks", 19)) strcpy(h->metric, "FMKS");
  else if (!strncmp(tr, "mks", 19)) strcpy(h->metric, "MKS");
  else rc = -2; /* eks and other transforms: no device metric */
  rc |= par_get(par, "parthenon/mesh", "nx1", 'i', &h->N1, 0);
  rc |= par_get(par, "parthenon/mesh", "nx2", 'i', &h->N2, 0);
  rc |= par_get(par, "parthenon/mesh", "nx3", 'i', &h->N3, 0);
  rc |= par_get(par, "parthenon/mesh", "x1min", 'd', &h->startx[1], 0);
  rc |= par_get(par, "parthenon/mesh", "x1max", 'd', &x1max, 0);
  rc |= par_get(par, "parthenon/mesh", "x2min", 'd', &h->startx[2], 0);
  rc |= par_get(par, "parthenon/mesh", "x2max", 'd', &x2max, 0);
  rc |= par_get(par, "parthenon/mesh", "x3min", 'd', &h->startx[3], 0);
  rc |= par_get(par, "parthenon/mesh", "x3max", 'd', &x3max, 0);
  par_get(par, "parthenon/mesh", "nghost", 'i', &ki->nghost, 0);
  h->gam = 1.444444444444444; /* model/kharma/model.c:53 */
  par_get(par, "GRMHD", "gamma", 'd', &h->gam, 0);
  rc |= par_get(par, "coordinates", "a", 'd', &h->a, 0);
  rc |= par_get(par, "coordinates", "r_in", 'd', &h->Rin, 0);
  rc |= par_get(par, "coordinates", "r_out", 'd', &h->Rout, 0);
  if (rc == 0) {
    rc |= par_get(par, "coordinates", "hslope", 'd', &h->hslope, 0);
    if (!strcmp(h->metric, "FMKS")) {
      rc |= par_get(par, "coordinates", "mks_smooth", 'd', &h->mks_smooth, 0);
      rc |= par_get(par, "coordinates", "poly_xt", 'd', &h->poly_xt, 0);
      rc |= par_get(par, "coordinates", "poly_alpha", 'd', &h->poly_alpha, 0);
    }
  }
  ki->game = 1.333333333333333;
  ki->gamp = 1.666666666666667;
  if (par_get(par, "electrons", "on", 's', el, sizeof(el)) == 0 && !strncmp(el, "true", 19)) {
    ki->has_electrons = 1;
    par_get(par, "electrons", "gamma_e", 'd', &ki->game, 0);
    par_get(par, "electrons", "gamma_p", 'd', &ki->gamp, 0);
  }
  free(par);
  if (rc) return rc < 0 ? rc : -1;
  h->n_prim = 8;
  h->dx[1] = (x1max - h->startx[1]) / h->N1;
  h->dx[2] = (x2max - h->startx[2]) / h->N2;
  h->dx[3] = (x3max - h->startx[3]) / h->N3;
  if (ki->num_meshblocks < 1 || ki->mb[0] < 1 || ki->mb[1] < 1 || ki->mb[2] < 1) return -1;
  return 0;
}

/* Read a dataset whole as doubles; its element count must be n. */
static int r_all(hid_t f, const char *name, size_t n, double *out)
{
  if (H5Lexists(f, name, H5P_DEFAULT) <= 0) return -1;
  hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
  if (ds < 0) return -1;
  hid_t sp = H5Dget_space(ds);
  int rc = ((size_t)H5Sget_simple_extent_npoints(sp) == n) ? 0 : -1;
  if (rc == 0 && H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) rc = -1;
  H5Sclose(sp);
  H5Dclose(ds);
  return rc;
}

/* load_kharma_data's mesh assembly (model/kharma/model.c:1267-1460): the
 * meshblocks of prims.rho, prims.u, prims.uvec and prims.B placed by
 * Blocks/loc.lx123, into [N1][N2][N3][8] (rho u U1 U2 U3 B1 B2 B3).
 * Format "new": vectors [block][component][x3][x2][x1]; "old" (KHARMA v5):
 * [block][x3][x2][x1][component], scalars with a trailing axis of 1. */
int ipole_h5_read_kharma_prims(const char *path, const ipole_dump_header *h, const ipole_kharma_info *ki,
                               double *prims)
{
  const int nmb = ki->num_meshblocks, n1 = ki->mb[0], n2 = ki->mb[1], n3 = ki->mb[2];
  const int old = !strcmp(ki->format, "old");
  if (!old && strcmp(ki->format, "new")) return -1;
  const size_t per = (size_t)n1 * n2 * n3, ns = (size_t)nmb * per;
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  int *loc = malloc(sizeof(int) * 3 * (size_t)nmb);
  double *rho = malloc(sizeof(double) * ns), *uu = malloc(sizeof(double) * ns);
  double *uv = malloc(sizeof(double) * 3 * ns), *bv = malloc(sizeof(double) * 3 * ns);
  int rc = (loc && rho && uu && uv && bv) ? 0 : -1;
  if (rc == 0) {
    hid_t ds = H5Lexists(f, "Blocks/loc.lx123", H5P_DEFAULT) > 0 ? H5Dopen2(f, "Blocks/loc.lx123", H5P_DEFAULT) : -1;
    if (ds < 0) rc = -1;
    else {
      hid_t sp = H5Dget_space(ds);
      if ((size_t)H5Sget_simple_extent_npoints(sp) != 3 * (size_t)nmb
          || H5Dread(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, loc) < 0)
        rc = -1;
      H5Sclose(sp);
      H5Dclose(ds);
    }
  }
  if (rc == 0) rc = r_all(f, "prims.rho", ns, rho) | r_all(f, "prims.u", ns, uu) | r_all(f, "prims.uvec", 3 * ns, uv)
                   | r_all(f, "prims.B", 3 * ns, bv);
  H5Fclose(f);
  for (int b = 0; b < nmb && rc == 0; b++) {
    const int *L = loc + 3 * b;
    if (L[0] < 0 || L[1] < 0 || L[2] < 0 || (L[0] + 1) * n1 > h->N1 || (L[1] + 1) * n2 > h->N2
        || (L[2] + 1) * n3 > h->N3)
      rc = -1;
    for (int kb = 0; kb < n3 && rc == 0; kb++)
      for (int jb = 0; jb < n2; jb++)
        for (int ib = 0; ib < n1; ib++) {
          const size_t s = (size_t)b * per + ((size_t)kb * n2 + jb) * n1 + ib;
          const size_t i = (size_t)L[0] * n1 + ib, j = (size_t)L[1] * n2 + jb, k = (size_t)L[2] * n3 + kb;
          double *o = prims + ((i * h->N2 + j) * h->N3 + k) * 8;
          o[0] = rho[s];
          o[1] = uu[s];
          for (int c = 0; c < 3; c++) {
            const size_t v = old ? s * 3 + c : ((size_t)b * 3 + c) * per + ((size_t)kb * n2 + jb) * n1 + ib;
            o[2 + c] = uv[v];
            o[5 + c] = bv[v];
          }
        }
  }
  free(loc);
  free(rho);
  free(uu);
  free(uv);
  free(bv);
  return rc;
}

static int w_attr(hid_t loc, const char *name, hid_t type, int n, const void *v)
{
  hsize_t d = (hsize_t)n;
  hid_t sp = n == 1 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, &d, NULL);
  hid_t a = H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
  int rc = (a < 0 || H5Awrite(a, type, v) < 0) ? -1 : 0;
  if (a >= 0) H5Aclose(a);
  H5Sclose(sp);
  return rc;
}

/* A KHARMA (Parthenon) file of the layout load_kharma_data reads, for
 * synthetic inputs: the parameter file text par, the meshblocks in the
 * order of order[] (block b holds the block at grid position order[b] in
 * x1-fastest numbering; NULL: in that order), primitives [N1][N2][N3][8]. */
int ipole_h5_write_kharma(const char *path, const ipole_dump_header *h, const ipole_kharma_info *ki,
                          const char *par, const int *order, const double *prims)
{
  const int n1 = ki->mb[0], n2 = ki->mb[1], n3 = ki->mb[2];
  if (h->N1 % n1 || h->N2 % n2 || h->N3 % n3) return -1;
  const int b1 = h->N1 / n1, b2 = h->N2 / n2, b3 = h->N3 / n3, nmb = b1 * b2 * b3;
  const int old = !strcmp(ki->format, "old");
  const size_t per = (size_t)n1 * n2 * n3, ns = (size_t)nmb * per;
  int *loc = malloc(sizeof(int) * 3 * (size_t)nmb);
  double *rho = malloc(sizeof(double) * ns), *uu = malloc(sizeof(double) * ns);
  double *uv = malloc(sizeof(double) * 3 * ns), *bv = malloc(sizeof(double) * 3 * ns);
  if (!loc || !rho || !uu || !uv || !bv) return -1;
  for (int b = 0; b < nmb; b++) {
    const int g = order ? order[b] : b;
    loc[3 * b] = g % b1;
    loc[3 * b + 1] = (g / b1) % b2;
    loc[3 * b + 2] = g / (b1 * b2);
    for (int kb = 0; kb < n3; kb++)
      for (int jb = 0; jb < n2; jb++)
        for (int ib = 0; ib < n1; ib++) {
          const size_t s = (size_t)b * per + ((size_t)kb * n2 + jb) * n1 + ib;
          const size_t i = (size_t)loc[3 * b] * n1 + ib, j = (size_t)loc[3 * b + 1] * n2 + jb,
                       k = (size_t)loc[3 * b + 2] * n3 + kb;
          const double *o = prims + ((i * h->N2 + j) * h->N3 + k) * 8;
          rho[s] = o[0];
          uu[s] = o[1];
          for (int c = 0; c < 3; c++) {
            const size_t v = old ? s * 3 + c : ((size_t)b * 3 + c) * per + ((size_t)kb * n2 + jb) * n1 + ib;
            uv[v] = o[2 + c];
            bv[v] = o[5 + c];
          }
        }
  }
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  int rc = f < 0 ? -1 : 0;
  if (rc == 0) {
    hid_t info = H5Gcreate2(f, "Info", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    rc |= w_attr(info, "NumMeshBlocks", H5T_STD_I32LE, 1, &nmb);
    rc |= w_attr(info, "MeshBlockSize", H5T_STD_I32LE, 3, ki->mb);
    rc |= w_attr(info, "Time", H5T_IEEE_F64LE, 1, &h->t);
    H5Gclose(info);
    hid_t in = H5Gcreate2(f, "Input", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hid_t st = H5Tcopy(H5T_C_S1);
    H5Tset_size(st, strlen(par) + 1);
    rc |= w_attr(in, "File", st, 1, par);
    H5Tclose(st);
    H5Gclose(in);
    hid_t bl = H5Gcreate2(f, "Blocks", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hsize_t dl[2] = {(hsize_t)nmb, 3};
    rc |= w_array(bl, "loc.lx123", 2, dl, H5T_STD_I32LE, loc);
    H5Gclose(bl);
    hsize_t ds[5] = {(hsize_t)nmb, (hsize_t)n3, (hsize_t)n2, (hsize_t)n1, 1};
    hsize_t dv_new[5] = {(hsize_t)nmb, 3, (hsize_t)n3, (hsize_t)n2, (hsize_t)n1};
    hsize_t dv_old[5] = {(hsize_t)nmb, (hsize_t)n3, (hsize_t)n2, (hsize_t)n1, 3};
    rc |= w_array(f, "prims.rho", old ? 5 : 4, ds, H5T_IEEE_F64LE, rho);
    rc |= w_array(f, "prims.u", old ? 5 : 4, ds, H5T_IEEE_F64LE, uu);
    rc |= w_array(f, "prims.uvec", 5, old ? dv_old : dv_new, H5T_IEEE_F64LE, uv);
    rc |= w_array(f, "prims.B", 5, old ? dv_old : dv_new, H5T_IEEE_F64LE, bv);
    H5Fclose(f);
  }
  free(loc);
  free(rho);
  free(uu);
  free(uv);
  free(bv);
  return rc;
}

/* /fluid_header of an image made from a KHARMA dump (output_hdf5,
 * model/kharma/model.c:1954-2121): the numeric keys of the grid and
 * coordinates, and for fmks the geom/fmks and geom/mmks groups. */
int ipole_h5_write_kharma_fluid_header(const char *image_path, const ipole_dump_header *h,
                                       const ipole_kharma_info *ki)
{
  hid_t f = H5Fopen(image_path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) return -1;
  hid_t fh = H5Gcreate2(f, "fluid_header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = fh < 0 ? -1 : 0;
  if (rc == 0) {
    const double r_eh = 1. + sqrt(1. - h->a * h->a);
    const double x1max = h->startx[1] + h->N1 * h->dx[1], x2max = h->startx[2] + h->N2 * h->dx[2],
                 x3max = h->startx[3] + h->N3 * h->dx[3];
    rc |= w_d(fh, "a", h->a);
    rc |= w_d(fh, "gam", h->gam);
    rc |= w_d(fh, "game", ki->game);
    rc |= w_d(fh, "gamp", ki->gamp);
    rc |= w_d(fh, "gamma", h->gam);
    rc |= w_d(fh, "hslope", h->hslope);
    rc |= w_d(fh, "mks_smooth", h->mks_smooth);
    rc |= w_i(fh, "nghost", ki->nghost);
    rc |= w_i(fh, "nx1", h->N1);
    rc |= w_i(fh, "nx2", h->N2);
    rc |= w_i(fh, "nx3", h->N3);
    rc |= w_i(fh, "nx1_mb", ki->mb[0]);
    rc |= w_i(fh, "nx2_mb", ki->mb[1]);
    rc |= w_i(fh, "nx3_mb", ki->mb[2]);
    rc |= w_d(fh, "poly_alpha", h->poly_alpha);
    rc |= w_d(fh, "poly_xt", h->poly_xt);
    rc |= w_d(fh, "r_eh", r_eh);
    rc |= w_d(fh, "r_in", h->Rin);
    rc |= w_d(fh, "r_out", h->Rout);
    rc |= w_d(fh, "rhor", r_eh);
    rc |= w_d(fh, "startx1", h->startx[1]);
    rc |= w_d(fh, "startx2", h->startx[2]);
    rc |= w_d(fh, "startx3", h->startx[3]);
    rc |= w_d(fh, "t", h->t);
    rc |= w_d(fh, "x1min", h->startx[1]);
    rc |= w_d(fh, "x1max", x1max);
    rc |= w_d(fh, "x2min", h->startx[2]);
    rc |= w_d(fh, "x2max", x2max);
    rc |= w_d(fh, "x3min", h->startx[3]);
    rc |= w_d(fh, "x3max", x3max);
    hid_t geom = H5Gcreate2(fh, "geom", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    rc |= w_d(geom, "startx1", h->startx[1]);
    rc |= w_d(geom, "startx2", h->startx[2]);
    rc |= w_d(geom, "startx3", h->startx[3]);
    if (!strcmp(h->metric, "FMKS")) {
      const char *names[2] = {"fmks", "mmks"};
      for (int q = 0; q < 2; q++) {
        hid_t g = H5Gcreate2(geom, names[q], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        rc |= w_d(g, "a", h->a);
        rc |= w_d(g, "hslope", h->hslope);
        rc |= w_d(g, "mks_smooth", h->mks_smooth);
        rc |= w_d(g, "poly_alpha", h->poly_alpha);
        rc |= w_d(g, "poly_xt", h->poly_xt);
        rc |= w_d(g, "r_eh", r_eh);
        rc |= w_d(g, "r_in", h->Rin);
        rc |= w_d(g, "r_out", h->Rout);
        H5Gclose(g);
      }
    }
    H5Gclose(geom);
    H5Gclose(fh);
  }
  H5Fclose(f);
  return rc ? -1 : 0;
}
