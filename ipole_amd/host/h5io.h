/* HDF5 I/O of the host (dumps and image files); see h5io.c. */
#ifndef IPOLE_H5IO_H
#define IPOLE_H5IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  char metric[32];
  int N1, N2, N3, n_prim;
  double gam, a, hslope, Rin, Rout, poly_xt, poly_alpha, mks_smooth;
  double startx[4], dx[4];
  double t, dump_cadence;
} ipole_dump_header;

typedef struct {
  char version[32], githash[32];
  int nx, ny, only_unpolarized, qu_conv, model_kind, reverse_field;
  double freqcgs, scale, dsource;
  double dx, dy, fovx_dsource, fovy_dsource, rcam, thetacam, phicam, rotcam, fovx, fovy, xoff, yoff;
  double Xcam[4];
  double t, sigma_cut, trat_small, trat_large, beta_crit;
  double L_unit, M_unit, T_unit, Thetae_unit, Mdot, MdotEdd, Ladv;
} ipole_image_meta;

int ipole_h5_write_dump(const char *path, const ipole_dump_header *h, const double *prims);
int ipole_h5_read_dump_header(const char *path, ipole_dump_header *h);
int ipole_h5_read_dump_prims(const char *path, const ipole_dump_header *h, double *prims);
int ipole_h5_write_image(const char *path, const ipole_image_meta *mt, const double *image,
                         const double *imageS, const double *taus);
int ipole_h5_copy_fluid_header(const char *dump_path, const char *image_path);
int ipole_h5_read_image(const char *path, int nx, int ny, double *image, double *imageS,
                        double *taus, double *Ftot, double *Ftot_unpol);

/* Trace file, src/io.c:279-543: open appends to an existing file, else
 * creates it with the image header and zero-filled datasets of nx x ny
 * pixels and maxnstep steps (returns a file handle, or -1).  write stores
 * pixel (i, j): its nstep and records 0 .. nstep of IPGPU_TRACE_NF doubles
 * as ipgpu_trace_pixel returns them. */
int64_t ipole_h5_trace_open(const char *path, const ipole_image_meta *mt, int nx, int ny, int maxnstep);
int ipole_h5_trace_write(int64_t fid, int i, int j, int nstep, const double *rec);
int ipole_h5_trace_close(int64_t fid);

/* KHARMA (Parthenon) dumps, model/kharma/model.c:407-1010, 1251-1470.  The
 * header reader fills the dump header from the parameter file stored in
 * the file (grid, coordinates, adiabatic index) and the Info attributes;
 * format ("new", or "old" for KHARMA v5 files) is read from *ki, the rest
 * of *ki is filled.  Returns 0, -1 (unreadable), or -2 for a coordinate
 * transform other than fmks / mks.  The primitives come back as
 * ipole_h5_read_dump_prims returns an iharm dump's: [N1][N2][N3][8].  The
 * writer makes files of that layout for synthetic inputs (order[b]: grid
 * position, x1 fastest, of stored block b; NULL for in order). */
typedef struct {
  char format[8];
  int num_meshblocks, mb[3], nghost, has_electrons;
  double game, gamp;
} ipole_kharma_info;
int ipole_h5_read_kharma_header(const char *path, ipole_dump_header *h, ipole_kharma_info *ki);
int ipole_h5_read_kharma_prims(const char *path, const ipole_dump_header *h, const ipole_kharma_info *ki,
                               double *prims);
int ipole_h5_write_kharma(const char *path, const ipole_dump_header *h, const ipole_kharma_info *ki,
                          const char *par, const int *order, const double *prims);
/* /fluid_header of an image of a KHARMA dump (model/kharma/model.c:1954-2121):
 * its numeric keys. */
int ipole_h5_write_kharma_fluid_header(const char *image_path, const ipole_dump_header *h,
                                       const ipole_kharma_info *ki);

/* Any dataset: its rank and dims (up to 16), or -1; its values as doubles. */
int ipole_h5_dataset_shape(const char *path, const char *name, int64_t *dims);
int ipole_h5_read_doubles(const char *path, const char *name, double *out);

#ifdef __cplusplus
}
#endif
#endif
