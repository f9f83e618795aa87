/* HDF5 I/O of the host (dumps and image files); see h5io.c. */
#ifndef IPOLE_H5IO_H
#define IPOLE_H5IO_H

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

#ifdef __cplusplus
}
#endif
#endif
