/*
 * CPU build of the GPU library's per-ray code (ipole_amd/csrc/ipg_pipeline.h)
 * for checking the kernel arithmetic against the oracle in a container
 * without a GPU.  Test infrastructure only: it is not reachable from the
 * library's C-ABI and is never used as a fallback.
 */
#include <cstring>
#include <vector>

#include "ipg_pipeline.h"

using namespace ipg;

/* The FP32 copy of the cell records that the library's mixed-precision
 * batch transfer gathers (k_dump_f32 on the device). */
static std::vector<float> dump_f32(const ipgpu_geom &G, const ipgpu_model &M, const double *dump)
{
  std::vector<float> o;
  if (M.kind != IPGPU_MODEL_IHARM || !dump) return o;
  const size_t nc = (size_t)(G.N1 + 2) * (G.N2 + 2) * (G.N3 + 2);
  o.assign(nc * CELL32, 0.f);
  for (size_t i = 0; i < nc; i++)
    for (int f = 0; f < 10; f++) o[i * CELL32 + f] = (float)dump[i * CELL_STRIDE + f];
  return o;
}

extern "C" {

/* metric + derivatives + inverse at (X1, X2): 42 doubles (r th g gc d1 d2 sqrtmg) */
void emu_metric(const ipgpu_geom *G, double X1, double X2, double *out)
{
  Metric m;
  metric(*G, X1, X2, m);
  out[0] = m.r;
  out[1] = m.th;
  memcpy(out + 2, m.g, 10 * sizeof(double));
  memcpy(out + 12, m.gc, 10 * sizeof(double));
  memcpy(out + 22, m.d1, 10 * sizeof(double));
  memcpy(out + 32, m.d2, 10 * sizeof(double));
  out[42] = m.sqrtmg;
}

void emu_accel(const ipgpu_geom *G, const double *X, const double *K, double *acc, double *M16)
{
  Metric m;
  metric(*G, X[1], X[2], m);
  geodesic_accel(m, K, acc);
  double M[4][4];
  transport_matrix(m, K, M);
  memcpy(M16, M, 16 * sizeof(double));
}

void emu_bessel(double x, double *k3, double *emx)
{
  bessel_k012(x, k3, *emx);
}

/* fluid state + coefficients at X, K: Ucon Ucov Bcon Bcov ne thetae B (19),
 * ju ku (2), pol coefficients (11) */
void emu_point(const ipgpu_geom *G, const ipgpu_model *M, const double *dump, int emission_type,
               const double *X, const double *K, double *out)
{
  Metric m;
  metric(*G, X[1], X[2], m);
  Fluid F;
  fluid(*G, *M, dump, X, m, F);
  memcpy(out, F.Ucon, 4 * sizeof(double));
  memcpy(out + 4, F.Ucov, 4 * sizeof(double));
  memcpy(out + 8, F.Bcon, 4 * sizeof(double));
  memcpy(out + 12, F.Bcov, 4 * sizeof(double));
  out[16] = F.ne;
  out[17] = F.thetae;
  out[18] = F.B;
  Coef c;
  coefficients(F, K, emission_type, true, out[19], out[20], c);
  double *o = out + 21;
  o[0] = c.jI; o[1] = c.jQ; o[2] = c.jU; o[3] = c.jV;
  o[4] = c.aI; o[5] = c.aQ; o[6] = c.aU; o[7] = c.aV;
  o[8] = c.rQ; o[9] = c.rU; o[10] = c.rV;
}

/* The pipeline's elementary functions: kind 0 exp_, 1 sincos_, 2 sincospi_,
 * 3 rcp, 4 log_, 5 cbrt_, 6 pow_(x, -8/3); out holds one value (two for the
 * sincos forms). */
void emu_math(int kind, const double *x, int n, double *out)
{
  for (int i = 0; i < n; i++) {
    double *o = out + 2 * i;
    o[1] = 0.;
    if (kind == 0) o[0] = exp_(x[i]);
    else if (kind == 1) sincos_(x[i], o[0], o[1]);
    else if (kind == 2) sincospi_(x[i], o[0], o[1]);
    else if (kind == 3) o[0] = rcp(x[i]);
    else if (kind == 4) o[0] = log_(x[i]);
    else if (kind == 5) o[0] = cbrt_(x[i]);
    else o[0] = pow_(x[i], -8. / 3.);
  }
}

/* The affine form of the plasma-frame Stokes step: S_out = T S_in + cv. */
void emu_stokes_affine(const double *c11, double x, const double *S_in, double *S_out)
{
  Coef c;
  c.jI = c11[0]; c.jQ = c11[1]; c.jU = c11[2]; c.jV = c11[3];
  c.aI = c11[4]; c.aQ = c11[5]; c.aU = c11[6]; c.aV = c11[7];
  c.rQ = c11[8]; c.rU = c11[9]; c.rV = c11[10];
  double T[4][4], cv[4];
  stokes_affine(c, x, T, cv);
  for (int i = 0; i < 4; i++)
    S_out[i] = cv[i] + T[i][0] * S_in[0] + T[i][1] * S_in[1] + T[i][2] * S_in[2] + T[i][3] * S_in[3];
}

/* The same with the mixed-precision step (stokes_affine_f32). */
void emu_stokes_affine_f32(const double *c11, double x, const double *S_in, double *S_out)
{
  Coef c;
  c.jI = c11[0]; c.jQ = c11[1]; c.jU = c11[2]; c.jV = c11[3];
  c.aI = c11[4]; c.aQ = c11[5]; c.aU = c11[6]; c.aV = c11[7];
  c.rQ = c11[8]; c.rU = c11[9]; c.rV = c11[10];
  double T[4][4], cv[4];
  stokes_affine_f32(c, x, T, cv);
  for (int i = 0; i < 4; i++)
    S_out[i] = cv[i] + T[i][0] * S_in[0] + T[i][1] * S_in[1] + T[i][2] * S_in[2] + T[i][3] * S_in[3];
}

int emu_plasma_tetrad(const ipgpu_geom *G, const double *X, const double *U, const double *K,
                      const double *B, double *E16)
{
  Metric m;
  metric(*G, X[1], X[2], m);
  double E[4][4];
  int fl = plasma_tetrad(U, K, B, m.g, m.sqrtmg, E);
  memcpy(E16, E, 16 * sizeof(double));
  return fl;
}

/* One image, per pixel (i*ny + j): out7, nstep, flag.  cap = maxnstep. */
int emu_render(const ipgpu_geom *G, const ipgpu_model *M, const ipgpu_image *P, const double *dump,
               const int *pix, int npix, double *out7, int *nstep, int *flag)
{
  int status = 0;
  const std::vector<float> d32 = P->transfer_fp32 ? dump_f32(*G, *M, dump) : std::vector<float>();
  const float *d32p = d32.empty() ? nullptr : d32.data();
#pragma omp parallel
  {
    std::vector<double> rec((size_t)P->maxnstep * REC);
    std::vector<int> nt(P->target_nturns >= 0 ? P->maxnstep : 1);
    const bool pol = !P->only_unpolarized;
    const int stride = pol ? PK : PK_UNPOL;
    std::vector<double> pkt;
#pragma omp for schedule(dynamic, 1)
    for (int n = 0; n < npix; n++) {
      int i = pix[2 * n], j = pix[2 * n + 1];
      int *ntp = P->target_nturns >= 0 ? nt.data() : nullptr;
      int ns = trace_ray(*G, *P, i, j, P->maxnstep, rec.data(), ntp);
      nstep[n] = ns;
      if (ns >= P->maxnstep - 1) {
#pragma omp critical
        status = -10;
        continue;
      }
      /* the same per-step packets and per-ray chain as the GPU kernels */
      pkt.assign((size_t)(ns > 0 ? ns : 1) * stride, 0.);
      double js, ks;
      if (P->transfer_fp32) {
        for (int q = 0; q < ns; q++)
          prep_step<true>(*G, *M, *P, ns - q, rec.data(), ntp, dump, pkt.data() + (size_t)q * stride, pol, unit_scale(),
                          nullptr, NoVG(), d32p);
        far_end_coefficients<true>(*G, *M, *P, ns, rec.data(), dump, js, ks, unit_scale());
      } else {
        for (int q = 0; q < ns; q++)
          prep_step<false>(*G, *M, *P, ns - q, rec.data(), ntp, dump, pkt.data() + (size_t)q * stride, pol, unit_scale());
        far_end_coefficients<false>(*G, *M, *P, ns, rec.data(), dump, js, ks, unit_scale());
      }
      int nrad;
      flag[n] = chain_ray(*G, *P, ns, pkt.data(), stride, js, ks, out7 + 7 * (size_t)n, nrad);
    }
  }
  return status;
}

/* The batch path on the CPU: geodesics traced with the image parameters
 * Ptrace, transfer of each ray fused (no packet memory) for the job Pjob
 * (whose frequency may differ: K and dl rescaled) with n_e scaled by ne_s
 * and B by its square root. */
int emu_render_scaled(const ipgpu_geom *G, const ipgpu_model *M, const ipgpu_image *Ptrace,
                      const ipgpu_image *Pjob, const double *dump, const int *pix, int npix, double ne_s,
                      double *out7, int *nstep, int *flag)
{
  int status = 0;
  Scale sc;
  sc.kf = Pjob->freq / Ptrace->freq;
  sc.ikf = 1. / sc.kf;
  sc.ne = ne_s;
  sc.b = sqrt(ne_s);
  const std::vector<float> d32 = Pjob->transfer_fp32 ? dump_f32(*G, *M, dump) : std::vector<float>();
  const float *d32p = d32.empty() ? nullptr : d32.data();
#pragma omp parallel
  {
    std::vector<double> rec((size_t)Ptrace->maxnstep * REC);
    std::vector<int> nt(Ptrace->target_nturns >= 0 ? Ptrace->maxnstep : 1);
#pragma omp for schedule(dynamic, 1)
    for (int n = 0; n < npix; n++) {
      int i = pix[2 * n], j = pix[2 * n + 1];
      int *ntp = Ptrace->target_nturns >= 0 ? nt.data() : nullptr;
      int ns = trace_ray(*G, *Ptrace, i, j, Ptrace->maxnstep, rec.data(), ntp);
      nstep[n] = ns;
      if (ns >= Pjob->maxnstep - 1) {
#pragma omp critical
        status = -10;
        continue;
      }
      int nrad;
      flag[n] = Pjob->transfer_fp32
          ? transfer_fused<true>(*G, *M, *Pjob, ns, rec.data(), ntp, dump, sc, out7 + 7 * (size_t)n, nrad, nullptr,
                                 d32p)
          : transfer_fused<false>(*G, *M, *Pjob, ns, rec.data(), ntp, dump, sc, out7 + 7 * (size_t)n, nrad);
    }
  }
  return status;
}

/* The multi-job transfer (transfer_fused_multi) of nj <= MJ jobs over the
 * rays traced with Ptrace: out7 (npix x nj x 7), flag (npix x nj). */
int emu_render_multi(const ipgpu_geom *G, const ipgpu_model *M, const ipgpu_image *Ptrace, int nj,
                     const ipgpu_image *Pjobs, const double *ne_s, const double *dump, const int *pix, int npix,
                     double *out7, int *nstep, int *flag)
{
  if (nj < 1 || nj > MJ) return -1;
  JobCfg cfg[MJ];
  const ipgpu_image *Pp[MJ];
  for (int j = 0; j < nj; j++) {
    cfg[j].sc.kf = Pjobs[j].freq / Ptrace->freq;
    cfg[j].sc.ikf = 1. / cfg[j].sc.kf;
    cfg[j].sc.ne = ne_s[j];
    cfg[j].sc.b = sqrt(ne_s[j]);
    cfg[j].etype = Pjobs[j].emission_type;
    cfg[j].pol = !Pjobs[j].only_unpolarized;
    cfg[j].floors = Pjobs[j].stokes_floors;
    cfg[j].cj = Pjobs[j].counterjet;
    Pp[j] = &Pjobs[j];
  }
  const bool f32 = Pjobs[0].transfer_fp32 != 0;
  const std::vector<float> d32 = f32 ? dump_f32(*G, *M, dump) : std::vector<float>();
  const float *d32p = d32.empty() ? nullptr : d32.data();
  int status = 0;
#pragma omp parallel
  {
    std::vector<double> rec((size_t)Ptrace->maxnstep * REC);
    std::vector<int> nt(Ptrace->target_nturns >= 0 ? Ptrace->maxnstep : 1);
    double mi[16], nst[16 * MJ];
#pragma omp for schedule(dynamic, 1)
    for (int n = 0; n < npix; n++) {
      int i = pix[2 * n], j = pix[2 * n + 1];
      int *ntp = Ptrace->target_nturns >= 0 ? nt.data() : nullptr;
      int ns = trace_ray(*G, *Ptrace, i, j, Ptrace->maxnstep, rec.data(), ntp);
      nstep[n] = ns;
      if (ns >= Ptrace->maxnstep - 1) {
#pragma omp critical
        status = -10;
        continue;
      }
      int nrad[MJ];
      if (f32)
        transfer_fused_multi<true>(*G, *M, nj, cfg, Pp, ns, rec.data(), ntp, dump, out7 + 7 * (size_t)nj * n,
                                   flag + (size_t)nj * n, nrad, mi, nst, 1, d32p);
      else
        transfer_fused_multi<false>(*G, *M, nj, cfg, Pp, ns, rec.data(), ntp, dump, out7 + 7 * (size_t)nj * n,
                                    flag + (size_t)nj * n, nrad, mi, nst, 1);
    }
  }
  return status;
}

/* Loss and gradients of the L2 image loss against target (npix x 5, the
 * reference layout) over all pixels of the image: out[0..3] = loss,
 * dL/dln(M_unit), dL/dR_high, dL/dR_low; gcell (ghosted cells, zeroed by
 * the caller) receives dL/d(n_e scale) per cell; Simg (npix x 4). */
int emu_adjoint(const ipgpu_geom *G, const ipgpu_model *M, const ipgpu_image *P, const double *dump,
                const double *target, const double *w, double *out, double *gcell, double *Simg, int seg_len)
{
  const int npix = P->nx * P->ny;
  int status = 0;
  double tot[4] = {0., 0., 0., 0.};
#pragma omp parallel
  {
    std::vector<double> rec((size_t)P->maxnstep * REC);
    std::vector<double> pkt, aux;
    std::vector<Herm> ckpt, seg(CK);
    double loc[4] = {0., 0., 0., 0.};
#pragma omp for schedule(dynamic, 1)
    for (int n = 0; n < npix; n++) {
      int i = n / P->ny, j = n % P->ny;
      int ns = trace_ray(*G, *P, i, j, P->maxnstep, rec.data(), nullptr);
      if (ns >= P->maxnstep - 1) {
#pragma omp critical
        status = -10;
        continue;
      }
      pkt.assign((size_t)(ns > 0 ? ns : 1) * PK, 0.);
      aux.assign((size_t)(ns > 0 ? ns : 1) * AX, 0.);
      for (int q = 0; q < ns; q++)
        prep_step<false>(*G, *M, *P, ns - q, rec.data(), nullptr, dump, pkt.data() + (size_t)q * PK, true,
                         unit_scale(), aux.data() + (size_t)q * AX);
      double o[4] = {0., 0., 0., 0.};
      const double tg[4] = {target[5 * n], target[5 * n + 1], target[5 * n + 2], target[5 * n + 3]};
      if (seg_len <= 0) {
        ckpt.resize(ns / CK + 1);
        adjoint_ray(*G, *P, ns, pkt.data(), aux.data(), tg, w, nullptr, dump, ckpt.data(), seg.data(), gcell, o,
                    Simg + 4 * (size_t)n);
      } else {
        /* the GPU's segmented reverse pass: maps of segments of seg_len
         * packets (the first one's source column only), composition, then
         * each segment's reverse pass */
        const int nseg = (ns + seg_len - 1) / seg_len;
        std::vector<double> map((size_t)(nseg > 0 ? nseg : 1) * ASEG_MAP, 0.);
        for (int k = 0; k < nseg; k++) {
          const int q0 = k * seg_len, q1 = q0 + seg_len < ns ? q0 + seg_len : ns;
          for (int b = (k == 0 ? NPAR : 0); b <= NPAR; b++) {
            Herm N;
            segment_column<1>(pkt.data(), q0, q1, b, N);
            for (int c = 0; c < NPAR; c++) map[(size_t)k * ASEG_MAP + (size_t)b * NPAR + c] = herm_par(N, c);
          }
        }
        std::vector<Herm> Ns(nseg > 0 ? nseg : 1), Ls(nseg > 0 ? nseg : 1);
        o[0] = adjoint_compose(*G, *P, nseg, map.data(), tg, w, nullptr, Ns.data(), Ls.data(), Simg + 4 * (size_t)n);
        ckpt.resize(seg_len / CK + 1);
        for (int k = 0; k < nseg; k++) {
          const int q0 = k * seg_len, q1 = q0 + seg_len < ns ? q0 + seg_len : ns;
          double g3[3] = {0., 0., 0.};
          adjoint_range(*G, *P, pkt.data(), aux.data(), q0, q1, Ns[k], Ls[k], dump, ckpt.data(), seg.data(), gcell,
                        g3);
          for (int q = 0; q < 3; q++) o[q + 1] += g3[q];
        }
      }
      for (int q = 0; q < 4; q++) loc[q] += o[q];
    }
#pragma omp critical
    for (int q = 0; q < 4; q++) tot[q] += loc[q];
  }
  for (int q = 0; q < 4; q++) out[q] = tot[q];
  return status;
}

/* Every ray through the segmented chain (composed segment maps of seg_len
 * steps; on the GPU only rays longer than the segment threshold take this
 * path). */
int emu_render_segmented(const ipgpu_geom *G, const ipgpu_model *M, const ipgpu_image *P, const double *dump,
                         const int *pix, int npix, int seg_len, double *out7, int *nstep, int *flag)
{
  int status = 0;
#pragma omp parallel
  {
    std::vector<double> rec((size_t)P->maxnstep * REC);
    std::vector<double> pkt, segmap;
    std::vector<int> segflag;
#pragma omp for schedule(dynamic, 1)
    for (int n = 0; n < npix; n++) {
      int i = pix[2 * n], j = pix[2 * n + 1];
      int ns = trace_ray(*G, *P, i, j, P->maxnstep, rec.data(), nullptr);
      nstep[n] = ns;
      if (ns >= P->maxnstep - 1) {
#pragma omp critical
        status = -10;
        continue;
      }
      pkt.assign((size_t)(ns > 0 ? ns : 1) * PK, 0.);
      for (int q = 0; q < ns; q++)
        prep_step<false>(*G, *M, *P, ns - q, rec.data(), nullptr, dump, pkt.data() + (size_t)q * PK, true,
                         unit_scale());
      const int nseg = (ns + seg_len - 1) / seg_len;
      segmap.assign((size_t)(nseg > 0 ? nseg : 1) * SEGCOL * NPAR, 0.);
      segflag.assign(nseg > 0 ? nseg : 1, 0);
      for (int sg = 0; sg < nseg; sg++) {
        int q0, q1;
        segment_bounds(ns, seg_len, sg, q0, q1);
        for (int b = 0; b < SEGCOL; b++) {
          double *o = segmap.data() + ((size_t)sg * SEGCOL + b) * NPAR;
          if (b == SEG_I) {
            segment_intensity(pkt.data(), q0, q1, o);
            continue;
          }
          Herm N;
          int fl = segment_column(pkt.data(), q0, q1, b, N);
          for (int p = 0; p < NPAR; p++) o[p] = herm_par(N, p);
          if (b == NPAR) segflag[sg] = fl;
        }
      }
      double js, ks;
      far_end_coefficients<false>(*G, *M, *P, ns, rec.data(), dump, js, ks, unit_scale());
      int nrad;
      flag[n] = chain_ray_segmented(*G, *P, ns, seg_len, segmap.data(), segflag.data(), js, ks,
                                    out7 + 7 * (size_t)n, nrad);
    }
  }
  return status;
}

/* Trajectory of one pixel: nstep+1 records of REC doubles. */
int emu_trace(const ipgpu_geom *G, const ipgpu_image *P, int i, int j, double *rec)
{
  return trace_ray(*G, *P, i, j, P->maxnstep, rec, nullptr);
}

/* The same trajectory traced as the two-phase render does: at most budget
 * steps, then resumed from the last record.  *paused = 1 if the first pass
 * stopped at the budget. */
int emu_trace_two_pass(const ipgpu_geom *G, const ipgpu_image *P, int i, int j, int budget, double *rec,
                       int *paused)
{
  int fin = 1;
  int ns = trace_ray(*G, *P, i, j, P->maxnstep, rec, nullptr, 0, budget, &fin);
  *paused = !fin;
  if (fin) return ns;
  return trace_ray(*G, *P, 0, 0, P->maxnstep, rec, nullptr, ns);
}

/* Trace output of one pixel as k_trace_pixel forms it: out receives
 * maxnstep records of IPGPU_TRACE_NF doubles. */
int emu_trace_pixel(const ipgpu_geom *G, const ipgpu_model *M, const ipgpu_image *P, const double *dump,
                    int prims_ok, int i, int j, double *out)
{
  const int cap = P->maxnstep;
  std::vector<double> rec((size_t)cap * REC);
  std::vector<int> nt(cap, 0);
  memset(out, 0, sizeof(double) * (size_t)cap * IPGPU_TRACE_NF);
  int ns = 0;
  trace_pixel_records(*G, *M, *P, i, j, cap, prims_ok, dump, rec.data(), nt.data(), out, &ns);
  return ns;
}

}
