/*
 * ipgpu: C-ABI of the MI355X (gfx950) per-pixel polarized ray pipeline.
 *
 * This library replaces the OpenMP pixel loop of the reference ipole
 * (src/main.c:536-615, get_pixel src/main.c:822-883, save_pixel :885-910)
 * with GPU kernels, behind plain C entry points: pointers, sizes and POD
 * structs, no torch or HIP types.  One context drives one GPU and one HIP
 * stream; calls on one context come from one host thread.
 *
 * Status codes: 0 on success, negative on failure (never exit()); the
 * message is available from ipgpu_last_error().  IPGPU_ERR_MAXNSTEP is the
 * condition on which the reference calls exit(-10) (src/main.c:841-845).
 */
#ifndef IPGPU_H
#define IPGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPGPU_OK 0
#define IPGPU_ERR_ARG (-1)
#define IPGPU_ERR_HIP (-2)
#define IPGPU_ERR_NAN (-3)
#define IPGPU_ERR_NOMEM (-4)
#define IPGPU_ERR_STATE (-5)
#define IPGPU_ERR_MAXNSTEP (-10)

/* coordinate systems, same codes as the reference (src/coordinates.h:450-464) */
#define IPGPU_METRIC_MKS 0
#define IPGPU_METRIC_FMKS 2
#define IPGPU_METRIC_EMINKOWSKI 5

/* fluid models */
#define IPGPU_MODEL_IHARM 1   /* model/iharm/model.c: ghosted dump fields */
#define IPGPU_MODEL_SPHERE 2  /* model/sphere/model.c: isothermal sphere */

/* number of ghosted per-cell fields the iharm model reads on the path:
 * U1 U2 U3 B1 B2 B3 (primitives) ne thetae b sigma */
#define IPGPU_NFIELDS 10

/* Coordinate globals of src/coordinates.h:25-33 as the model sets them
 * (model/iharm/model.c:979-1065, :326). */
typedef struct ipgpu_geom {
  int metric;
  int N1, N2, N3;
  double a, hslope, poly_xt, poly_alpha, mks_smooth, poly_norm;
  double startx[4], stopx[4], dx[4], cstartx[4], cstopx[4];
  double Rin, Rout, Rh, rmin_geo, rmax_geo;
} ipgpu_geom;

/* Model constants (model/iharm/model.c:49-94, model/sphere/model.c:20-96). */
typedef struct ipgpu_model {
  int kind;
  double L_unit;                         /* cm per GM/c^2 */
  double th_beg, sigma_cut, sigma_cut_high;
  double sph_R0, sph_Ne0, sph_Thetae0, sph_B0;
  /* sphere sigma and beta inside R0 (model/sphere/model.c:172-188), read
   * only by the trace output: Thetae0 tp_over_te / beta0 and beta0 */
  double sph_sigma, sph_beta;
} ipgpu_model;

/* Units and electron model for the on-device dump derivation
 * (model/iharm/model.c:595-608, :610-717; ELECTRONS = 2). */
typedef struct ipgpu_units {
  double M_unit, MBH_solar;
  double trat_small, trat_large, beta_crit;
  double gam, game, gamp;
} ipgpu_units;

/* One image: the Params fields get_pixel reads (src/par.h:14-73) after the
 * camera set-up of src/main.c:126-174. */
typedef struct ipgpu_image {
  int nx, ny;
  double Xcam[4];            /* native_coord(rcam, thetacam, phicam) */
  double fovx, fovy;         /* DX / rcam, DY / rcam */
  double freq;               /* freqcgs * HPL / (ME CL^2) */
  double freqcgs;
  double rotcam;             /* radians */
  double xoff, yoff;
  double eps;
  int maxnstep;
  int old_centering;
  int only_unpolarized;
  int qu_conv;
  int stokes_floors;
  int target_nturns;
  int subring_dtheta;
  int counterjet;
  int emission_type;         /* 4 (Dexter thermal, default) or 1 (Pandya thermal) */
  int transfer_fp32;         /* 1: the coefficients' shape functions and the plasma-frame
                                Stokes step (rotation, absorption/emission solution, its
                                matrix and source terms) in FP32, and the fluid gather
                                reads the cell records rounded to FP32 (a copy the
                                library writes from the dump before each render or batch
                                transfer; not in ipgpu_render_noalloc, which allocates
                                nothing and reads the FP64 records); geodesics,
                                connection, interpolation, tetrads and N in FP64.  No
                                counterpart in the reference, which is FP64 throughout */
} ipgpu_image;

typedef struct ipgpu_ctx ipgpu_ctx;

/* Context on GPU `device`; replaces nothing in the reference (set-up). */
int ipgpu_create(ipgpu_ctx **ctx, int device);
void ipgpu_destroy(ipgpu_ctx *ctx);
const char *ipgpu_last_error(const ipgpu_ctx *ctx);

/* Geometry and model constants; mirror the globals init_model leaves behind
 * (model/iharm/model.c:295-389). */
int ipgpu_set_geometry(ipgpu_ctx *ctx, const ipgpu_geom *geom);
int ipgpu_set_model(ipgpu_ctx *ctx, const ipgpu_model *model);

/* Upload the model's ghosted (N1+2)(N2+2)(N3+2) arrays, x3 fastest, as the
 * iharm plugin holds them after load (data[0]->p[U1..B3], ne, thetae, b,
 * sigma: model/iharm/model.c:104-114, utils.c:41-98).  The caller keeps
 * ownership; this is a copy into the device layout. */
int ipgpu_upload_fields(ipgpu_ctx *ctx, const double *const fields[IPGPU_NFIELDS],
                        size_t ncells_ghosted);

/* Upload dump primitives [N1][N2][N3][nprim] (rho u U1 U2 U3 B1 B2 B3 ...)
 * and derive the fields on the device: b from the primitives
 * (model/iharm/model.c:1954-2010), ghost zones (:1314-1370), n_e, Theta_e,
 * sigma (:610-717). */
int ipgpu_upload_prims(ipgpu_ctx *ctx, const double *prims, int nprim, const ipgpu_units *units);

/* Copy the derived ghosted fields back (11 arrays: the 10 above and beta),
 * for checking the derivation. */
int ipgpu_download_fields(ipgpu_ctx *ctx, double *out11, size_t ncells_ghosted);

/* Render one image: the fast-light pixel loop of src/main.c:536-615.
 * Outputs use the reference layout (i*ny + j; imageS has 5 entries per
 * pixel: I Q U V tauF), already multiplied by freqcgs^3 as save_pixel does.
 * Any output pointer may be NULL: results then stay on the device and can
 * be fetched with ipgpu_fetch().  nstep/flag are per pixel (nstep of the
 * trajectory, oddflag bits of integrate_emission).  imageS is ignored with
 * only_unpolarized. */
int ipgpu_render(ipgpu_ctx *ctx, const ipgpu_image *img, double *image, double *imageS,
                 double *taus, int *nstep, int *flag);
int ipgpu_fetch(ipgpu_ctx *ctx, double *image, double *imageS, double *taus, int *nstep,
                int *flag);

/* ipgpu_render for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops and harnesses that forbid allocation).
 * No counterpart in the reference.  The outputs are the caller's DEVICE
 * arrays (image, taus, nstep, flag: nx*ny entries; imageS: nx*ny*5; any may
 * be NULL), in ipgpu_render's layout and units; the inputs are the
 * context's geometry, model and uploaded dump.  The call allocates nothing:
 * its working memory is one static device array of IPGPU_SCRATCH_BYTES per
 * GPU (held for as long as the process runs), shared by every call on that
 * GPU and left as the last call used it, so calls that overlap on one GPU
 * (from several contexts or threads) race.  The rays are traced once to
 * count their steps and again, in as many chunks as the scratch needs, to
 * store their records, and each ray's transfer runs in one lane
 * (transfer_fused, the arithmetic of ipgpu_render_batch): results agree
 * with ipgpu_render to roundoff.  Returns IPGPU_ERR_ARG when the image's
 * per-ray tables and two rays of maxnstep records exceed the scratch (about
 * 85 million rays at maxnstep 50000), or for target_nturns >= 0. */
#define IPGPU_SCRATCH_BYTES ((size_t)1 << 30)
int ipgpu_render_noalloc(ipgpu_ctx *ctx, const ipgpu_image *img, double *image, double *imageS, double *taus,
                         int *nstep, int *flag);

/* Adaptive image refinement (src/main.c:505-792; keys src/par.c:77-84,
 * 172-177): the image's nx x ny is the finest level, nx_min x ny_min the
 * coarsest; levels = log2(nx / nx_min) + 1 on the even parts of nx, nx_min,
 * and an even side gains one internal row or column (the 2^N + 1 scheme).
 * The coarsest grid is traced, then each finer level traces the pixels
 * whose interpolation error exceeds both refine_abs (relative to the
 * estimated mean intensity) and refine_rel, where the lower neighbour is
 * brighter than refine_cut, and interpolates the others (linearly, or
 * from the nearest traced neighbour).  The internal grid must be
 * (nx_min - 1) 2^(levels - 1) + 1 on both axes.  nx_min <= 0 or
 * nx_min == nx renders as ipgpu_render.  Outputs: the nx x ny image as
 * ipgpu_render returns it (interpolated pixels have nstep = flag = 0) and
 * interp (may be NULL): 1 where a pixel was interpolated.  image must not
 * be NULL; the other outputs may be. */
typedef struct ipgpu_refine {
  int nx_min, ny_min;
  double refine_abs, refine_rel, refine_cut;
  int nearest_neighbor;
} ipgpu_refine;
#define IPGPU_MAX_LEVELS 16
int ipgpu_render_adaptive(ipgpu_ctx *ctx, const ipgpu_image *img, const ipgpu_refine *rf, double *image,
                          double *imageS, double *taus, int *nstep, int *flag, int *interp);

/* After ipgpu_render_adaptive: the number of levels, the estimated mean
 * intensity of the coarse pass (Iavg, src/main.c:617-632), and per level
 * (index 0 the coarse pass) the pixels traced and the interpolated pixels
 * of the internal grid so far (the count src/main.c:776-786 prints).
 * traced and interpolated hold IPGPU_MAX_LEVELS entries; any may be NULL. */
int ipgpu_adaptive_stats(const ipgpu_ctx *ctx, int *levels, double *Iavg, long long *traced,
                         long long *interpolated);

/* One job of a batched render: an image (camera, size, frequency and
 * options as for ipgpu_render) of the uploaded dump, with n_e scaled by
 * m_unit_scale and B by its square root, i.e. the dump's M_unit times
 * m_unit_scale (1 for the uploaded one).  Output pointers (host memory, any
 * may be NULL) receive the job's image in the layout of ipgpu_render. */
typedef struct ipgpu_job {
  ipgpu_image image;
  double m_unit_scale;
  double *image_out, *imageS_out, *taus_out;
  int *nstep_out, *flag_out;
} ipgpu_job;

/* Batched render: many cameras, frequencies and M_unit values of one dump
 * in one pass (the reference runs one ipole process per image; the
 * fast-light pixel loop of src/main.c:536-615 per job).  Jobs with the same
 * camera, image size and integration options share one geodesic trace at
 * the frequency of the first of them (the geodesic in X does not depend on
 * the frequency).  A first trace counts every ray's steps and, when a slot
 * of records per ray fits the batch memory, keeps its first records there,
 * so that only the longer rays are continued into exact regions; otherwise
 * the rays are traced again, longest first, into exact slices of one
 * compact record buffer, in as many passes as the memory requires.  The
 * transfer of every (job, ray) pair then runs in one fused launch, longest
 * rays first.  Results agree with per-job ipgpu_render to
 * the roundoff of the rescaling.  On IPGPU_ERR_MAXNSTEP the job and pixel
 * are reported by ipgpu_bad_job.  Measurement: ipgpu_kernel_times gives the
 * batch's geodesic, transfer and total device time. */
int ipgpu_render_batch(ipgpu_ctx *ctx, int njobs, const ipgpu_job *jobs);
int ipgpu_bad_job(const ipgpu_ctx *ctx, int *job, int *i, int *j);

/* Device memory a batch may use for geodesic records (default 160 GiB):
 * records kept in place from the first trace when a slot per ray plus the
 * longer rays' regions fit, else exact slices in as many passes as this
 * requires. */
int ipgpu_set_batch_memory(ipgpu_ctx *ctx, size_t traj_bytes);

/* Render the last batch again with new m_unit_scale values, reusing its
 * geodesic records: the transfer alone runs (a captured HIP graph of the
 * job-table upload, status clears and transfer launch, replayed on later
 * calls).  jobs must hold the same images, in the same order, as the last
 * ipgpu_render_batch, whose records must have stayed in place (one pass; it
 * returns IPGPU_ERR_STATE after a batch in passes, or after ipgpu_render,
 * ipgpu_render_adaptive, ipgpu_adjoint, ipgpu_backward, ipgpu_trace_pixel
 * or ipgpu_set_geometry, which reuse or invalidate the records).  A new
 * dump of the same grid (ipgpu_upload_prims, ipgpu_scale_density) or a new
 * model is allowed: the geodesics depend on the geometry and cameras only.
 * The M_unit search of the reference's users, one ipole run per trial
 * M_unit (src/main.c:536-615 per run), becomes one trace and one transfer
 * per trial. */
int ipgpu_rerender_batch(ipgpu_ctx *ctx, int njobs, const ipgpu_job *jobs);

/* Batch launch options (a negative value keeps the current one): the
 * record slot per ray of the first trace (default 2048), the transfer as a
 * persistent work-queue launch (1, default) or one wave per 64 items (0),
 * the persistent waves (0: four per compute unit), and whether
 * ipgpu_rerender_batch replays a captured graph (1, default) or launches
 * eagerly (0).  No counterpart in the reference. */
int ipgpu_set_batch_options(ipgpu_ctx *ctx, int slot, int queue, int queue_waves, int rerender_graph);

/* Pick the batch options for this job set by measurement and keep them in
 * the context: the record slot (1024, 2048 or 4096; whole batches timed on
 * the device), then the transfer launch (one wave per 64 items, or the work
 * queue at one or two times the waves the transfer kernel keeps resident;
 * the transfer alone timed on the records left in place).  The last batch
 * stays resident for ipgpu_rerender_batch.  chosen (optional): slot, queue,
 * queue waves and resident waves per compute unit.  Output pointers of the
 * jobs are ignored.  No counterpart in the reference, whose only launch
 * choice is OpenMP's dynamic schedule over pixels (src/main.c:536). */
int ipgpu_autotune_batch(ipgpu_ctx *ctx, int njobs, const ipgpu_job *jobs, int chosen[4]);

/* The last batch's transfer work, for measurement: out[0] work items (one
 * lane each: a ray with up to three jobs of its trace group, whose geometry
 * they share, or with IPGPU_BATCH_MULTI=0 one job), out[1] the items'
 * transfer steps, out[2] their radiating steps (an item's counted once),
 * out[3] (job, ray) pairs, out[4] their transfer steps.  No counterpart in
 * the reference. */
int ipgpu_batch_counts(const ipgpu_ctx *ctx, long long out[5]);

/* Device memory a single-image render (ipgpu_render) may use for its ray
 * records and step packets (default 60% of the device's memory, and never
 * more than the device has free).  An image whose rays need more is
 * rendered in chunks of rays (whole 8x8 pixel tiles) that fit; the image
 * does not depend on the chunking.  The reference needs one trajectory of
 * maxnstep records per thread (src/main.c:835); this bounds the GPU's
 * equivalent, which holds the records of every ray of a chunk at once. */
int ipgpu_set_render_memory(ipgpu_ctx *ctx, size_t bytes);

/* How the single-image render's step-packet kernels read the dump cells of
 * the trilinear interpolation (src/grid.c:15-44, interp_scalar): on = 1
 * stages in LDS, per group of 8 consecutive steps of a ray, the box of
 * cells that spans their interpolation corners (at most 24 cells, loaded
 * with direct-to-LDS loads ahead of the step's metric; larger boxes read
 * the dump directly); on = 0 (default) reads each step's 8 corner records
 * from the dump.  Both give the same interpolated fields; the polarized
 * images of the two settings agree to ~1e-11 relative (the compiler fuses
 * the step's multiply-adds differently in the two kernel instances), the
 * unpolarized images bit for bit.  No counterpart in the reference. */
int ipgpu_set_gather_staging(ipgpu_ctx *ctx, int on);

/* Work layout of the last ipgpu_render: chunks of rays it ran in, rays whose
 * trace outgrew the per-ray record capacity and continued in full-length
 * regions (the capacity follows the longest ray of the previous render), and
 * the device memory the context holds (dump, records, packets, images).
 * Measurement only; no counterpart in the reference. */
int ipgpu_render_info(const ipgpu_ctx *ctx, int *chunks, int *continued_rays, size_t *device_bytes);

/* Gradient of the image loss
 *   L = sum over pixels of sum_s w_s (S_s - target_s)^2,  s = I, Q, U, V,
 * (S in the layout and units of ipgpu_render's imageS) with respect to the
 * dump's M_unit, R_high (trat_large), R_low (trat_small) and a per-cell
 * scale of n_e, by reverse mode through the polarized transfer (one
 * hand-written adjoint kernel: forward recurrence with checkpoints every 32
 * steps, recomputed per segment on the way back).  No counterpart in the
 * reference.  Needs a dump from ipgpu_upload_prims (the electron model's
 * R_high/R_low dependence is taken from it).
 *   target:  nx*ny*5 doubles, ipgpu_render's imageS layout (tauF ignored)
 *   weights: w_I, w_Q, w_U, w_V (NULL: all 1)
 *   out4:    L, dL/dM_unit, dL/dR_high, dL/dR_low
 *   grad_density: N1*N2*N3 doubles (x3 fastest), dL/ds_c of n_e -> s_c n_e
 *            at s = 1 for every interior cell c (ghost-zone copies folded
 *            back), or NULL
 *   imageS:  the image the loss was taken on (NULL: not returned)
 * The Stokes floors (stokes_floors = 1) are not differentiated. */
int ipgpu_adjoint(ipgpu_ctx *ctx, const ipgpu_image *img, const double *target, const double *weights,
                  double out4[4], double *grad_density, double *imageS);

/* Vector-Jacobian product of the polarized render: given dL/dS for the
 * image ipgpu_render would return (dL_dimageS: nx*ny*5 doubles in imageS
 * layout; the tauF entries are ignored), the gradient of L with respect to
 * the dump's M_unit, R_high and R_low (out3) and, if grad_density is not
 * NULL, to a per-cell scale of n_e at s = 1 (N1*N2*N3 doubles, x3 fastest,
 * ghost copies folded back), by the same reverse pass as ipgpu_adjoint.
 * imageS (may be NULL) receives the image.  No counterpart in the
 * reference; the same restrictions as ipgpu_adjoint. */
int ipgpu_backward(ipgpu_ctx *ctx, const ipgpu_image *img, const double *dL_dimageS, double out3[3],
                   double *grad_density, double *imageS);

/* Multiply n_e of every interior cell by scale[N1*N2*N3] (x3 fastest) and
 * refresh its ghost copies: the per-cell density scale whose gradient
 * ipgpu_adjoint returns.  No counterpart in the reference. */
int ipgpu_scale_density(ipgpu_ctx *ctx, const double *scale);

/* Trace output of one pixel: everything dump_var_along records
 * (src/io.c:279-543), the data of the reference's trace.h5 for pixel
 * (i, j).  out holds cap >= maxnstep records of IPGPU_TRACE_NF doubles;
 * record s has the values the reference stores at index s of its per-step
 * arrays (filled for s = 1 .. nstep - 1, and e2Transported also at 0; the
 * rest zero), at the offsets below.  rho, u and beta need the fields derived
 * on the device (ipgpu_upload_prims), else they are 0.  *nstep receives the
 * trajectory's step count.  One lane: a debugging and parity tool, not a
 * fast path. */
#define IPGPU_TRACE_NF 126
#define IPGPU_TR_PRIMS 0      /* 8: rho u U1 U2 U3 B1 B2 B3 */
#define IPGPU_TR_B 8
#define IPGPU_TR_NE 9
#define IPGPU_TR_THETAE 10
#define IPGPU_TR_NU 11
#define IPGPU_TR_BK_ANGLE 12
#define IPGPU_TR_SIGMA 13
#define IPGPU_TR_BETA 14
#define IPGPU_TR_DL 15        /* cgs */
#define IPGPU_TR_R 16
#define IPGPU_TR_TH 17
#define IPGPU_TR_PHI 18
#define IPGPU_TR_J_UNPOL 19
#define IPGPU_TR_K_UNPOL 20
#define IPGPU_TR_I_UNPOL 21
#define IPGPU_TR_X 22         /* 4 */
#define IPGPU_TR_KCON 26      /* 4 */
#define IPGPU_TR_J_INV 30     /* 4 */
#define IPGPU_TR_ALPHA_INV 34 /* 4 */
#define IPGPU_TR_RHO_INV 38   /* 4, [0] = 0 */
#define IPGPU_TR_STOKES 42    /* 4 */
#define IPGPU_TR_E2C 46       /* 4: e2Constructed */
#define IPGPU_TR_E2C_COV 50   /* 4: e2ConstructedCov */
#define IPGPU_TR_E2T 54       /* 4: e2Transported */
#define IPGPU_TR_E2T_COV 58   /* 4: e2TransportedCov */
#define IPGPU_TR_NTETRAD 62   /* 16 */
#define IPGPU_TR_NCOORD 78    /* 16 */
#define IPGPU_TR_ECON 94      /* 16 */
#define IPGPU_TR_ECOV 110     /* 16 */
int ipgpu_trace_pixel(ipgpu_ctx *ctx, const ipgpu_image *img, int i, int j, int cap, int *nstep, double *out);

/* Pixel (i, j) that exceeded maxnstep in the last render (IPGPU_ERR_MAXNSTEP). */
int ipgpu_bad_pixel(const ipgpu_ctx *ctx, int *i, int *j);

/* Device time of the last render's kernels, from HIP events on the
 * context's stream: geodesic pass, transfer pass, total. */
int ipgpu_kernel_times(const ipgpu_ctx *ctx, double *ms_geodesic, double *ms_transfer,
                       double *ms_total);

/* Per-stage device time of the last render (HIP events).  Returns 0 after a
 * one-pass render: ms[0..3] = geodesic pass (k_trace), transfer step packets
 * (k_scan + k_prep), transfer recurrence (k_seg + k_chain), output layout
 * (k_finalize).  Returns 1 after a streamed render: ms[0] = the geodesic
 * launches (k_trace_budget, k_trace_resume_mask per later phase), ms[1] =
 * the span of the transfer of the rays finished in the bounded phases (on
 * their own streams, concurrent with later phases), ms[2] = from the end of
 * the last geodesic launch until the streams have joined (transfer of the
 * longest rays, the wait), ms[3] = k_finalize.  Measurement only; the
 * reference prints one wallclock figure (src/main.c:1038-1040). */
int ipgpu_stage_times(const ipgpu_ctx *ctx, double ms[4]);

/* Tuning of ipgpu_render (no counterpart in the reference, whose pixel loop
 * has none).  trace_budget, trace_phases: the streamed render integrates the
 * geodesics in trace_phases bounded phases of trace_budget j / trace_phases
 * steps (j = 1 .. trace_phases, at most 4), then to the end (defaults 3584
 * and 2; trace_budget 0 = one pass; streaming applies when target_nturns < 0
 * and the record capacity exceeds the budget).  seg_min_steps: polarized
 * rays longer than this have their transfer composed from segments (default
 * 3584).  seg_len: steps per segment (default 384).  record_steps:
 * geodesic records held per ray (default 8192; raised to the longest ray of
 * a render for the next, within half the render memory; a ray past it
 * continues in a full-length region of its own).  Results do not depend on
 * the phases or the record capacity; seg_min_steps and seg_len regroup the
 * transfer arithmetic of the rays selected (agreement to roundoff).  A
 * negative value leaves a setting unchanged.  The environment variables
 * IPGPU_TRACE_BUDGET, IPGPU_TRACE_PHASES, IPGPU_SEG_MIN_STEPS and
 * IPGPU_SEG_LEN set the defaults of a new context. */
int ipgpu_set_tuning(ipgpu_ctx *ctx, int trace_budget, int trace_phases, int seg_min_steps, int seg_len,
                     int record_steps);

/* Pick the streamed render's settings for this image by measurement and
 * keep them in the context: the segment length (256, 384, 512), the segment
 * threshold (the last phase's rays, or also the phase before's), then the
 * bounded phases' budget (3/4 or 5/4 of the current), one at a time, each
 * timed as the better of two renders on the device.  chosen (optional):
 * trace_budget, trace_phases, seg_min_steps, seg_len.  The images do not
 * depend on the budget; the segment settings regroup the transfer
 * arithmetic (agreement to roundoff).  No counterpart in the reference. */
int ipgpu_autotune_render(ipgpu_ctx *ctx, const ipgpu_image *img, int chosen[4]);

/* Work counts of the last render: geodesic steps summed over rays (the
 * nstep of trace_geodesic) and radiating steps of the transfer pass (steps
 * with radiating_region true, src/ipolarray.c:122). */
int ipgpu_step_counts(const ipgpu_ctx *ctx, long long *geodesic_steps, long long *radiating_steps);

/* The context's HIP stream (a hipStream_t), for callers that time or order
 * their own work against it. */
void *ipgpu_stream(ipgpu_ctx *ctx);

/* Host helpers of the camera set-up: native_coord/root_find
 * (src/coordinates.c:375-442). */
int ipgpu_native_coord(const ipgpu_geom *geom, double r, double th_deg, double phi_deg,
                       double X[4]);

#ifdef __cplusplus
}
#endif
#endif
