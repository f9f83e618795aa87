/*
 * Op-counting build of the GPU library's per-ray code (test and measurement
 * infrastructure).  Every double in ipg_physics.h / ipg_pipeline.h becomes
 * a counting number type, so running the pipeline over sample pixels on the
 * CPU counts the floating-point operations one unit of each kernel performs:
 *   +, -, *, /, sqrt                  1 each
 *   fma                               2
 *   exp, log, pow, sin, cos, acos, cbrt, tanh   20 each (one transcendental)
 * Comparisons, fabs, floor/rint, fmin/fmax and ldexp are not counted.  The
 * pipeline's own exp_/sincos_/rcp (ipg_math.h) are counted as one
 * transcendental / one division each (IPG_OPCOUNT), as the unit convention
 * prices a transcendental, not the instructions of one implementation.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ipgpu.h"

struct Ops {
  uint64_t add = 0, mul = 0, div = 0, sqrt_ = 0, fma_ = 0, trans = 0;
  double flops() const { return double(add + mul + div + sqrt_) + 2. * double(fma_) + 20. * double(trans); }
};
static Ops g_ops;

struct Cd {
  double v;
  constexpr Cd() : v(0.) {}
  constexpr Cd(double x) : v(x) {}
  explicit operator int() const { return (int)v; }
  explicit operator float() const { return (float)v; }
  explicit operator long long() const { return (long long)v; }
  Cd &operator+=(Cd o) { g_ops.add++; v += o.v; return *this; }
  Cd &operator-=(Cd o) { g_ops.add++; v -= o.v; return *this; }
  Cd &operator*=(Cd o) { g_ops.mul++; v *= o.v; return *this; }
  Cd &operator/=(Cd o) { g_ops.div++; v /= o.v; return *this; }
  Cd operator-() const { return Cd(-v); }
};
inline Cd operator+(Cd a, Cd b) { g_ops.add++; return Cd(a.v + b.v); }
inline Cd operator-(Cd a, Cd b) { g_ops.add++; return Cd(a.v - b.v); }
inline Cd operator*(Cd a, Cd b) { g_ops.mul++; return Cd(a.v * b.v); }
inline Cd operator/(Cd a, Cd b) { g_ops.div++; return Cd(a.v / b.v); }
inline bool operator<(Cd a, Cd b) { return a.v < b.v; }
inline bool operator>(Cd a, Cd b) { return a.v > b.v; }
inline bool operator<=(Cd a, Cd b) { return a.v <= b.v; }
inline bool operator>=(Cd a, Cd b) { return a.v >= b.v; }
inline bool operator==(Cd a, Cd b) { return a.v == b.v; }
inline bool operator!=(Cd a, Cd b) { return a.v != b.v; }
#define T1(name)                              \
  inline Cd name(Cd a)                        \
  {                                           \
    g_ops.trans++;                            \
    return Cd(std::name(a.v));                \
  }
T1(exp) T1(log) T1(sin) T1(cos) T1(acos) T1(cbrt) T1(tanh)
#undef T1
inline Cd pow(Cd a, Cd b) { g_ops.trans++; return Cd(std::pow(a.v, b.v)); }
inline Cd sqrt(Cd a) { g_ops.sqrt_++; return Cd(std::sqrt(a.v)); }
inline Cd fabs(Cd a) { return Cd(std::fabs(a.v)); }
inline Cd floor(Cd a) { return Cd(std::floor(a.v)); }
inline Cd rint(Cd a) { return Cd(std::rint(a.v)); }
inline Cd fmin(Cd a, Cd b) { return Cd(std::fmin(a.v, b.v)); }
inline Cd fmax(Cd a, Cd b) { return Cd(std::fmax(a.v, b.v)); }
inline Cd ldexp(Cd a, int e) { return Cd(std::ldexp(a.v, e)); }
inline Cd frexp(Cd a, int *e) { return Cd(std::frexp(a.v, e)); }
inline Cd fma(Cd a, Cd b, Cd c) { g_ops.fma_++; return Cd(std::fma(a.v, b.v, c.v)); }
inline Cd fmod(Cd a, Cd b) { g_ops.div++; return Cd(std::fmod(a.v, b.v)); }
inline bool isnan(Cd a) { return std::isnan(a.v); }
inline void sincos(Cd x, Cd *s, Cd *c)
{
  g_ops.trans += 2;
  s->v = std::sin(x.v);
  c->v = std::cos(x.v);
}

#define IPG_OPCOUNT 1
typedef double ipg_f64;
#define double Cd
#include "ipg_pipeline.h"
#undef double

using namespace ipg;

extern "C" {

/* Counts over the given pixels of one image (iharm dump as 16-double cell
 * records, or null for the sphere).  out[0..8]: geodesic steps, transfer
 * steps, radiating steps, flops of trace_ray, of prep_step on radiating
 * steps, of prep_step on non-radiating steps, of chain_ray, of
 * far_end_coefficients, number of rays, of transfer_fused (the batch's
 * FP64 fused transfer of the same ray), of segment_column over the ray's
 * 384-step segments (17 columns), of segment_intensity over them.
 * per_ray_multi (optional, 3 per pixel): transfer_fused_multi's flops for
 * the first 1, 2, 3 of three jobs (the image's frequency and M_unit, 230/86
 * of its frequency at 0.4 M_unit, 345/86 at 2.5 M_unit).
 * per_ray (optional, 3 per pixel):
 * the pixel's transfer steps, radiating steps and transfer_fused flops. */
int opcount_render(const ipgpu_geom *G, const ipgpu_model *M, const ipgpu_image *P, const double *dump,
                   const int *pix, int npix, double *out, double *per_ray, double *per_ray_multi)
{
  std::vector<Cd> rec((size_t)P->maxnstep * REC);
  std::vector<int> nt(P->target_nturns >= 0 ? P->maxnstep : 1);
  const bool pol = !P->only_unpolarized;
  const int stride = pol ? PK : PK_UNPOL;
  std::vector<Cd> pkt;
  const Cd *d = reinterpret_cast<const Cd *>(dump);
  for (int q = 0; q < 12; q++) out[q] = 0.;
  for (int n = 0; n < npix; n++) {
    int i = pix[2 * n], j = pix[2 * n + 1];
    int *ntp = P->target_nturns >= 0 ? nt.data() : nullptr;
    Ops o0 = g_ops;
    int ns = trace_ray(*G, *P, i, j, P->maxnstep, rec.data(), ntp);
    out[3] += g_ops.flops() - o0.flops();
    out[0] += ns;
    if (ns >= P->maxnstep - 1) return -10;
    pkt.assign((size_t)(ns > 0 ? ns : 1) * stride, Cd(0.));
    for (int q = 0; q < ns; q++) {
      Ops a = g_ops;
      prep_step(*G, *M, *P, ns - q, rec.data(), ntp, d, pkt.data() + (size_t)q * stride, pol, unit_scale());
      bool rad = ((int)pkt[(size_t)q * stride + PK_META]) & 1;
      double f = g_ops.flops() - a.flops();
      if (rad) {
        out[2] += 1;
        out[4] += f;
      } else {
        out[5] += f;
      }
    }
    out[1] += ns;
    Ops b = g_ops;
    Cd js, ks;
    far_end_coefficients(*G, *M, *P, ns, rec.data(), d, js, ks, unit_scale());
    out[7] += g_ops.flops() - b.flops();
    Ops c = g_ops;
    Cd o7[7];
    int nrad;
    chain_ray(*G, *P, ns, pkt.data(), stride, js, ks, o7, nrad);
    out[6] += g_ops.flops() - c.flops();
    out[8] += 1;
    Ops e = g_ops;
    Cd f7[7];
    int nrad_f;
    transfer_fused<false>(*G, *M, *P, ns, rec.data(), ntp, d, unit_scale(), f7, nrad_f);
    const double f_fused = g_ops.flops() - e.flops();
    out[9] += f_fused;
    if (pol) {
      const int nseg = (ns + SEG_LEN - 1) / SEG_LEN;
      Ops g = g_ops;
      for (int sg = 0; sg < nseg; sg++) {
        int q0, q1;
        segment_bounds(ns, SEG_LEN, sg, q0, q1);
        for (int b = 0; b <= NPAR; b++) {
          Herm N;
          segment_column<1>(pkt.data(), q0, q1, b, N);
        }
      }
      out[10] += g_ops.flops() - g.flops();
      Ops h = g_ops;
      for (int sg = 0; sg < nseg; sg++) {
        int q0, q1;
        segment_bounds(ns, SEG_LEN, sg, q0, q1);
        Cd u[NPAR];
        segment_intensity<1>(pkt.data(), q0, q1, u);
      }
      out[11] += g_ops.flops() - h.flops();
    }
    if (per_ray_multi) {
      ipgpu_image Pj[MJ];
      JobCfg cfg[MJ];
      const ipgpu_image *Pp[MJ];
      const double ratio[MJ] = {1., 230. / 86., 345. / 86.}, sne[MJ] = {1., 0.4, 2.5};
      for (int j = 0; j < MJ; j++) {
        Pj[j] = *P;
        Pj[j].freq = P->freq * ratio[j];
        Pj[j].freqcgs = P->freqcgs * ratio[j];
        cfg[j].sc.kf = ratio[j];
        cfg[j].sc.ikf = 1. / ratio[j];
        cfg[j].sc.ne = sne[j];
        cfg[j].sc.b = sqrt(sne[j]);
        cfg[j].etype = P->emission_type;
        cfg[j].pol = pol;
        cfg[j].floors = P->stokes_floors;
        cfg[j].cj = P->counterjet;
        Pp[j] = &Pj[j];
      }
      for (int nj = 1; nj <= MJ; nj++) {
        Cd mi[16], nst[16 * MJ], o7[7 * MJ];
        int fl[MJ], nr[MJ];
        Ops a = g_ops;
        transfer_fused_multi<false>(*G, *M, nj, cfg, Pp, ns, rec.data(), ntp, d, o7, fl, nr, mi, nst, 1);
        per_ray_multi[3 * n + nj - 1] = g_ops.flops() - a.flops();
      }
    }
    if (per_ray) {
      per_ray[3 * n] = ns;
      per_ray[3 * n + 1] = nrad_f;
      per_ray[3 * n + 2] = f_fused;
    }
  }
  return 0;
}
}
