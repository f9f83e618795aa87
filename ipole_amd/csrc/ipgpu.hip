/*
 * ipgpu: gfx950 kernels and the C-ABI of include/ipgpu.h.
 *
 * Pipeline of one image (replaces src/main.c:536-615):
 *   k_trace     one lane per ray: camera tetrad, initial wavevector and the
 *               backward geodesic (src/model_geodesics.c:38-171), writing one
 *               64-byte record per step (X1 X2 X3 K0 K1 K2 K3 dl) to HBM.
 *   k_scan      packet offsets: prefix sum of the per-ray step counts.
 *   k_prep      one lane per transfer step of every ray (src/ipolarray.c:
 *               59-247 minus the state): connection contractions for the
 *               parallel transport of N, fluid gathers from the dump in HBM,
 *               thermal coefficients and the plasma tetrad, written as one
 *               step packet; fully parallel over all steps of the image.
 *   k_chain     one lane per ray: the serial recurrence of N and I over its
 *               packets from the far end toward the camera, then project_N.
 *   k_finalize  reference output layout and units (save_pixel,
 *               src/main.c:885-910).
 * Rays are numbered in 8x8 pixel tiles so that one 64-lane wavefront traces
 * a compact patch of the image (similar path lengths, nearby dump cells).
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "ipg_physics.h"
#include "ipg_pipeline.h"
#include "ipgpu.h"

using namespace ipg;

namespace {

constexpr int TILE = 8;         /* pixels per tile side */
constexpr int WAVE = 64;
constexpr int NSTATUS = 4;      /* [0] overflowed rays [1] first maxnstep ray+1 [2] NaN pixel+1 */

/* The pixel a ray renders: with pix == nullptr the 8x8 tiles of the nx x ny
 * image (false for the padding rays of the edge tiles); otherwise
 * pix[ray] = (i << 16) | j, a pixel list (false for the padding entries, -1). */
__device__ __forceinline__ bool ray_pixel(const int *pix, int ray, int nx, int ny, int &i, int &j)
{
  if (pix) {
    const int v = pix[ray];
    i = v >> 16;
    j = v & 0xffff;
    return v >= 0;
  }
  int tiles_y = (ny + TILE - 1) / TILE;
  int t = ray / WAVE, l = ray % WAVE;
  int ti = t / tiles_y, tj = t % tiles_y;
  i = ti * TILE + l / TILE;
  j = tj * TILE + l % TILE;
  return i < nx && j < ny;
}

/* The records of a ray of a single-image render: its region of cap records
 * in traj, or, for a ray that outgrew it (rbase[ray] >= 0), its full-length
 * region at trajO + rbase[ray]. */
__device__ __forceinline__ const double *ray_rec(const double *traj, int cap, const long long *rbase,
                                                 const double *trajO, int ray)
{
  if (rbase) {
    const long long b = rbase[ray];
    if (b >= 0) return trajO + b;
  }
  return traj + (size_t)ray * cap * REC;
}

/* Issue priority of a kernel's waves among the waves of other kernels on
 * the same SIMD: the latency-bound kernels (the geodesic phases, the
 * chains, the segment maps of the rays traced last) take the issue slots
 * first and the packet kernels fill the rest. */
#define IPG_LATENCY_CRITICAL() __builtin_amdgcn_s_setprio(3)

/* pixel code of the status words: first bad pixel by atomicMin */
__device__ __forceinline__ int pixel_code(int i, int j) { return ((i << 16) | j) + 1; }

/* The two lanes (2k, 2k+1) of a lane pair exchange a value (DPP quad
 * permutation [1, 0, 3, 2] on both 32-bit halves). */
__device__ __forceinline__ double pair_swap(double v)
{
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, 0xB1, 0xF, 0xF, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi2, lo2);
}

/* Where trace_pair_t puts the records past its capacity: nowhere
 * (NoOverflow), or in chunks of OVF_CH records drawn from an arena
 * (OverflowChunks, the batch's first pass). */
struct NoOverflow {
  __device__ void store(int, bool, double, double, double, double) {}
};

constexpr int OVF_CH = 1024;   /* records per overflow chunk (64 KiB) */

/* A ray's records from step cap on, chunk by chunk: the pair's even lane
 * draws a chunk index from the arena's counter (nchunks of them), the odd
 * lane receives it, and table[c] keeps chunk c's index for the ray.  When a
 * chunk cannot be had (the arena used up, or more than maxc chunks) the ray
 * stores no more and ok is false; it is then continued as without chunks. */
struct OverflowChunks {
  double *arena;
  unsigned long long *counter;
  unsigned long long nchunks;
  int *table;
  int maxc, cap;
  double *cur;
  bool ok;
  __device__ __forceinline__ void store(int nstep, bool odd, double a, double b, double c, double d)
  {
    const int rel = nstep - cap;
    const int within = rel & (OVF_CH - 1);
    if (within == 0) {
      const int ci = rel / OVF_CH;
      double id = -1.;
      if (ok && ci < maxc && !odd) {
        const unsigned long long t = atomicAdd(counter, 1ull);
        id = t < nchunks ? (double)t : -1.;
      }
      const double io = pair_swap(id);
      id = odd ? io : id;
      if (!odd && ci < maxc) table[ci] = (int)id;
      if (id < 0.) ok = false;
      cur = ok ? arena + (size_t)(long long)id * OVF_CH * REC : nullptr;
    }
    if (ok) {
      double *o = cur + (size_t)within * REC + (odd ? 4 : 0);
      o[0] = a;
      o[1] = b;
      o[2] = c;
      o[3] = d;
    }
  }
};

/* trace_ray_t on a lane pair.  Both lanes hold the same X, K and take the
 * same decisions; the two metric evaluations of a step run one per lane
 * (the even lane at the half-step point, the odd lane at the step's end
 * point, which is the next step's start), and the accelerations each lane
 * forms from its metric are exchanged.  Half the instructions per lane of
 * the one-lane step, with the same operations on the same values, so the
 * records are bit-identical.  Each lane stores half of a record. */
template <int MK, class OV = NoOverflow>
__device__ __forceinline__ int trace_pair_t(const ipgpu_geom &G, const ipgpu_image &P, double X[4], double K[4],
                                            int nstep, int cap, double *rec, int nend, int *finished, bool odd,
                                            OV *ov = nullptr)
{
  Metric m;
  metric_k<MK>(G, X[1], X[2], m); /* both lanes: the start point */
  int done = 1;
  while (true) {
    /* r at X: the odd lane's metric (both lanes' on the first pass) */
    const double r_o = pair_swap(m.r);
    const double rX = odd ? m.r : r_o;
    if (stop_backward(G, rX, K)) break;
    if (nstep >= P.maxnstep - 1) break;
    if (nstep >= nend) {
      done = 0;
      break;
    }
    nstep++;
    double dl = stepsize(G, X, K, P.eps);
    double dlp = -dl;
    double a0[4];
    geodesic_accel(m, K, a0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const double o = pair_swap(a0[k]);
      a0[k] = odd ? a0[k] : o;
    }
    double Xh[4], Kh[4], Xn[4];
    half_step_acc(a0, X, K, dlp, Xh, Kh);
    for (int k = 0; k < 4; k++) Xn[k] = X[k] + dlp * Kh[k];
    metric_k<MK>(G, odd ? Xn[1] : Xh[1], odd ? Xn[2] : Xh[2], m);
    double acc[4];
    geodesic_accel(m, Kh, acc);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const double o = pair_swap(acc[k]);
      acc[k] = odd ? o : acc[k];
    }
    for (int k = 0; k < 4; k++) K[k] = K[k] - dlp * acc[k];
    for (int k = 0; k < 4; k++) X[k] = Xn[k];
    if (nstep < cap) {
      /* fields 0-3 from the even lane, 4-7 from the odd one */
      double *o = rec + (size_t)nstep * REC + (odd ? 4 : 0);
      o[0] = odd ? K[1] : X[1];
      o[1] = odd ? K[2] : X[2];
      o[2] = odd ? K[3] : X[3];
      o[3] = odd ? dl : K[0];
    } else if constexpr (!std::is_same<OV, NoOverflow>::value) {
      if (cap > 0) ov->store(nstep, odd, odd ? K[1] : X[1], odd ? K[2] : X[2], odd ? K[3] : X[3], odd ? dl : K[0]);
    }
  }
  if (finished) *finished = done;
  return nstep;
}

/* trace_ray on a lane pair: the camera start or the resumption is formed by
 * both lanes alike. */
template <class OV = NoOverflow>
__device__ __forceinline__ int trace_pair(const ipgpu_geom &G, const ipgpu_image &P, int pi, int pj, int cap,
                                          double *rec, int n0, int nend, int *finished, bool odd, OV *ov = nullptr)
{
  double X[4], K[4];
  if (n0 > 0) {
    Rec r;
    load_rec(rec, n0, r);
    X[0] = 0.;
    for (int k = 1; k < 4; k++) X[k] = r.X[k];
    for (int k = 0; k < 4; k++) K[k] = r.K[k];
  } else {
    double E[4][4];
    Metric m;
    for (int k = 0; k < 4; k++) X[k] = P.Xcam[k];
    camera_tetrad(G, X, P.old_centering, E, m);
    init_K(P, E, pi, pj, K);
    if (cap > 0 && !odd) store_rec(rec, X, K, 0.);
    n0 = 0;
  }
  switch (metric_kind(G)) {
  case MK_MINK: return trace_pair_t<MK_MINK, OV>(G, P, X, K, n0, cap, rec, nend, finished, odd, ov);
  case MK_MKS: return trace_pair_t<MK_MKS, OV>(G, P, X, K, n0, cap, rec, nend, finished, odd, ov);
  case MK_FMKS_INT: return trace_pair_t<MK_FMKS_INT, OV>(G, P, X, K, n0, cap, rec, nend, finished, odd, ov);
  default: return trace_pair_t<MK_FMKS, OV>(G, P, X, K, n0, cap, rec, nend, finished, odd, ov);
  }
}

/* trace_pair from the camera in two loops: to step cap - 1 storing the
 * slot's records, then, for a ray that goes on, the rest with its records
 * in arena chunks.  The second loop continues from the first one's X and K
 * (its start-point metric is the same function of the same X), so the
 * records are those of one loop; the first loop carries nothing of the
 * chunks' state. */
template <int MK>
__device__ __forceinline__ int trace_pair_slot_t(const ipgpu_geom &G, const ipgpu_image &P, double X[4], double K[4],
                                                 int cap, double *rec, bool odd, OverflowChunks &ov)
{
  int done = 1;
  const int ns = trace_pair_t<MK>(G, P, X, K, 0, cap, rec, cap - 1, &done, odd);
  if (done) return ns;
  return trace_pair_t<MK, OverflowChunks>(G, P, X, K, ns, cap, rec, 0x7fffffff, nullptr, odd, &ov);
}

__device__ __forceinline__ int trace_pair_slot(const ipgpu_geom &G, const ipgpu_image &P, int pi, int pj, int cap,
                                               double *rec, bool odd, OverflowChunks &ov)
{
  double X[4], K[4], E[4][4];
  Metric m;
  for (int k = 0; k < 4; k++) X[k] = P.Xcam[k];
  camera_tetrad(G, X, P.old_centering, E, m);
  init_K(P, E, pi, pj, K);
  if (!odd) store_rec(rec, X, K, 0.);
  switch (metric_kind(G)) {
  case MK_MINK: return trace_pair_slot_t<MK_MINK>(G, P, X, K, cap, rec, odd, ov);
  case MK_MKS: return trace_pair_slot_t<MK_MKS>(G, P, X, K, cap, rec, odd, ov);
  case MK_FMKS_INT: return trace_pair_slot_t<MK_FMKS_INT>(G, P, X, K, cap, rec, odd, ov);
  default: return trace_pair_slot_t<MK_FMKS>(G, P, X, K, cap, rec, odd, ov);
  }
}

/* the loss's component weights, by value */
struct AdjW {
  double w[4];
};

__global__ __launch_bounds__(WAVE) void k_trace(ipgpu_geom G, ipgpu_image P, const int *__restrict__ pix, int nrays, int cap,
                                                double *__restrict__ traj, int *__restrict__ nstep_out,
                                                int *__restrict__ nturns_buf, int *__restrict__ status,
                                                unsigned long long *__restrict__ counts)
{
  int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= nrays) return;
  int pi, pj;
  if (!ray_pixel(pix, ray, P.nx, P.ny, pi, pj)) {
    nstep_out[ray] = 0;
    return;
  }
  double *rec = traj + (size_t)ray * cap * REC;
  int *nt = nturns_buf ? nturns_buf + (size_t)ray * cap : nullptr;
  int nstep = trace_ray(G, P, pi, pj, cap, rec, nt);
  nstep_out[ray] = nstep;
  if (nstep >= cap) atomicAdd(&status[0], 1);
  atomicAdd(&counts[0], (unsigned long long)nstep);
}

/* k_trace without subring counts on lane pairs (trace_pair): the geodesics
 * of the one-pass render are then those of the streamed render, bit for
 * bit. */
__global__ __launch_bounds__(WAVE) void k_trace_pair(ipgpu_geom G, ipgpu_image P, const int *__restrict__ pix, int nrays, int cap,
                                                     double *__restrict__ traj, int *__restrict__ nstep_out,
                                                     int *__restrict__ status, unsigned long long *__restrict__ counts)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ray = t >> 1;
  const bool odd = t & 1;
  if (ray >= nrays) return;
  int pi, pj;
  if (!ray_pixel(pix, ray, P.nx, P.ny, pi, pj)) {
    if (!odd) nstep_out[ray] = 0;
    return;
  }
  int nstep = trace_pair(G, P, pi, pj, cap, traj + (size_t)ray * cap * REC, 0, 0x7fffffff, nullptr, odd);
  if (odd) return;
  nstep_out[ray] = nstep;
  if (nstep >= cap) atomicAdd(&status[0], 1);
  atomicAdd(&counts[0], (unsigned long long)nstep);
}

/* Exclusive prefix sum of the per-ray step counts (one workgroup): packet
 * offsets for k_prep / k_chain.  off has nrays + 1 entries. */
constexpr int SCAN_THREADS = 1024;
__global__ __launch_bounds__(SCAN_THREADS) void k_scan(int nrays, const int *__restrict__ nstep,
                                                       long long *__restrict__ off)
{
  __shared__ long long part[SCAN_THREADS];
  const int t = threadIdx.x;
  const int chunk = (nrays + SCAN_THREADS - 1) / SCAN_THREADS;
  const int lo = t * chunk, hi = min(lo + chunk, nrays);
  long long s = 0;
  for (int r = lo; r < hi; r++) s += nstep[r];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < SCAN_THREADS; d <<= 1) {
    long long v = (t >= d) ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long base = part[t] - s;
  for (int r = lo; r < hi; r++) {
    off[r] = base;
    base += nstep[r];
  }
  if (t == SCAN_THREADS - 1) off[nrays] = part[t];
}

/* Ray of the first step of each block of PREP_BLOCK packets: one thread per
 * ray writes the blocks that start inside its steps. */
constexpr int PREP_BLOCK = 256;
__global__ void k_blockmap(int nrays, const long long *__restrict__ off, int *__restrict__ blk_ray)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  long long b0 = (off[r] + PREP_BLOCK - 1) / PREP_BLOCK, b1 = (off[r + 1] + PREP_BLOCK - 1) / PREP_BLOCK;
  for (long long b = b0; b < b1; b++) blk_ray[b] = r;
}

/* LDS-staged dump gathers of the packet kernels.  In their waves each group
 * of 8 lanes holds 8 consecutive steps of one ray, whose interpolation
 * corners fall in a few neighbouring cells: the group stages the box of
 * cells that spans them (at most SG_CAP cells, else it gathers from the
 * dump as before) in LDS with direct-to-LDS loads, issued before the step's
 * metric is formed, and each lane then reads its 8 corners from LDS.  Cell
 * v of group g (box order: i, then j, then k fastest) is loaded by lane
 * 8 g + (v & 7); its 16-byte chunk ch (fields 2 ch, 2 ch + 1) lands in the
 * wave's 1 KiB row (v >> 3) * 5 + ch, at that lane's 16 bytes.  On the
 * 160x160 FMKS image ~80% of the steps' groups fit in 24 cells, which cuts
 * the corner-line requests of the gathers ~3x and turns them into
 * 80-byte runs. */
constexpr int SG_CAP = 24;                   /* cells per group box */
constexpr int SG_CPL = (SG_CAP + 7) / 8;     /* cells per lane */
constexpr int SG_ROWS = 5 * SG_CPL;          /* 1 KiB LDS rows per wave */
constexpr int SG_WAVE_DBL = SG_ROWS * 128;   /* doubles of LDS per wave */

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;

/* min / max over the 8 lanes of a group: DPP quad permutations [1,0,3,2]
 * and [2,3,0,1], then the half-row mirror (lane i <- lane 7 - i) */
__device__ __forceinline__ int group8_min(int x)
{
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, false));
  x = min(x, __builtin_amdgcn_update_dpp(x, x, 0x4E, 0xF, 0xF, false));
  return min(x, __builtin_amdgcn_update_dpp(x, x, 0x141, 0xF, 0xF, false));
}
__device__ __forceinline__ int group8_max(int x)
{
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x4E, 0xF, 0xF, false));
  return max(x, __builtin_amdgcn_update_dpp(x, x, 0x141, 0xF, 0xF, false));
}

/* A lane's view of its group's staged box; fetch() is prep_step's VG. */
struct StagedCells {
  const double *lds;   /* the wave's LDS rows */
  Cell c;              /* the lane's interpolation cell */
  int lo_i, lo_j, lo_k, bj, bk, slot0;
  bool use;            /* the group is staged and the lane needs the fields */
  Rec ti, tf;          /* the step's records, loaded before the staging loads
                        * (vmcnt is in order: a later record load would wait
                        * for them) */
  __device__ bool records(Rec &a, Rec &b) const
  {
    a = ti;
    b = tf;
    return true;
  }
  __device__ bool fetch(double v[10]) const
  {
    if (!use) return false;
    /* the direct-to-LDS loads count in vmcnt; nothing else reads them */
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    int off[8];
    for (int n = 0; n < 8; n++) {
      const int cv = ((c.i + (n >> 2) - lo_i) * bj + (c.j + ((n >> 1) & 1) - lo_j)) * bk + (c.k + (n & 1) - lo_k);
      off[n] = (cv >> 3) * 5 * 128 + 2 * (slot0 + (cv & 7));
    }
    interp10(c, [&](int n, int f) { return lds[off[n] + (f >> 1) * 128 + (f & 1)]; }, v);
    return true;
  }
};

/* Called by every lane of the wave (need = false for lanes without a step
 * or outside the dump); X1..X3 of the lane's interpolation point.  Forms
 * the group's box, issues its loads and fills S. */
__device__ __forceinline__ void stage_cells(const ipgpu_geom &G, const double *dump, bool need, double X1,
                                            double X2, double X3, double *lds_wave, StagedCells &S)
{
  const int lane = (int)(threadIdx.x & (WAVE - 1));
  const int t = lane & 7;
  if (need) locate(G, X1, X2, X3, S.c);
  const int BIG = 1 << 28;
  const int lo_i = group8_min(need ? S.c.i : BIG), hi_i = group8_max(need ? S.c.i : -BIG);
  const int lo_j = group8_min(need ? S.c.j : BIG), hi_j = group8_max(need ? S.c.j : -BIG);
  const int lo_k = group8_min(need ? S.c.k : BIG), hi_k = group8_max(need ? S.c.k : -BIG);
  const bool any = hi_i >= lo_i;
  const int bi = hi_i - lo_i + 2, bj = hi_j - lo_j + 2, bk = hi_k - lo_k + 2;
  const bool on = any && bi * bj <= SG_CAP && bi * bj * bk <= SG_CAP;
  const int V = on ? bi * bj * bk : 0;
  S.lds = lds_wave;
  S.lo_i = lo_i;
  S.lo_j = lo_j;
  S.lo_k = lo_k;
  S.bj = bj;
  S.bk = bk;
  S.slot0 = lane & ~7;
  S.use = on && need;
  const long s2 = G.N3 + 2, s1 = (long)(G.N2 + 2) * (G.N3 + 2);
  for (int m = 0; m < SG_CPL; m++) {
    const int v = m * 8 + t;
    if (v < V) {
      const int a = v / (bj * bk), r = v - a * (bj * bk), b = r / bk, kk = r - b * bk;
      const double *p = dump + CELL_STRIDE * ((long)(lo_i + a) * s1 + (long)(lo_j + b) * s2 + (lo_k + kk));
      for (int ch = 0; ch < 5; ch++)
        __builtin_amdgcn_global_load_lds((glb_void_t *)(p + 2 * ch), (lds_void_t *)(lds_wave + (m * 5 + ch) * 128),
                                         16, 0, 0);
    }
  }
}

/* stage_cells for step n of a ray's records (its interpolation point is
 * record n - 1, as in prep_step), whose records S keeps for prep_step;
 * act = false: the lane has no step. */
__device__ __forceinline__ void stage_step(const ipgpu_geom &G, const ipgpu_model &Mo, const double *dump, bool act,
                                           const double *rec, int n, const double *dummy, double *lds_wave,
                                           StagedCells &S)
{
  /* loaded on every path (a lane without a step reads a stand-in pair of
   * records, dummy: any two records that exist), so that no record load is
   * still counted when the staging loads are issued */
  load_rec(act ? rec : dummy, act ? n : 1, S.ti);
  load_rec(act ? rec : dummy, act ? n - 1 : 0, S.tf);
  /* each record value passes through an empty asm here, on every path, so
   * the record loads are waited for before the staging loads are issued
   * (a first use in a later branch would wait for those too) */
  for (int k = 1; k < 4; k++) {
    asm volatile("" : "+v"(S.ti.X[k]));
    asm volatile("" : "+v"(S.tf.X[k]));
  }
  for (int k = 0; k < 4; k++) {
    asm volatile("" : "+v"(S.ti.K[k]));
    asm volatile("" : "+v"(S.tf.K[k]));
  }
  asm volatile("" : "+v"(S.ti.dl));
  asm volatile("" : "+v"(S.tf.dl));
  const double X1 = S.tf.X[1], X2 = S.tf.X[2], X3 = S.tf.X[3];
  const bool need = act && Mo.kind != IPGPU_MODEL_SPHERE && in_domain(G, X1, X2);
  stage_cells(G, dump, need, X1, X2, X3, lds_wave, S);
}

/* The wave's LDS rows of a packet kernel's block */
__device__ __forceinline__ double *stage_rows(double *block_lds)
{
  return block_lds + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE)) * SG_WAVE_DBL;
}

/* One lane per transfer step of every ray: the state-independent part of
 * integrate_emission (transport matrices, fluid, coefficients, tetrad).
 * STAGE: the dump gathers staged in LDS as in the packet-list kernels. */
template <bool F32, bool STAGE>
__global__ __launch_bounds__(PREP_BLOCK) void k_prep(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P, int nrays,
                                                     int cap, int stride, long long total,
                                                     const long long *__restrict__ off,
                                                     const int *__restrict__ blk_ray,
                                                     const int *__restrict__ nstep_in,
                                                     const double *__restrict__ traj,
                                                     const int *__restrict__ nturns_buf,
                                                     const double *__restrict__ dump, double *__restrict__ pkt,
                                                     double *__restrict__ aux, const float *__restrict__ d32)
{
  __shared__ double sg_lds[STAGE ? (PREP_BLOCK / WAVE) * SG_WAVE_DBL : 1];
  long long g = (long long)blockIdx.x * PREP_BLOCK + threadIdx.x;
  bool act = g < total;
  /* ray = last r with off[r] <= g, from the block's first ray forward */
  int ray = blk_ray[blockIdx.x];
  if (act)
    while (ray + 1 < nrays && off[ray + 1] <= g) ray++;
  const int ns = nstep_in[ray];
  act = act && ns < P.maxnstep - 1;
  const int q = (int)(g - off[ray]);
  const double *rec = traj + (size_t)ray * cap * REC;
  const int *nt = nturns_buf ? nturns_buf + (size_t)ray * cap : nullptr;
  double *ax = aux ? aux + (size_t)g * AX : nullptr;
  if constexpr (STAGE) {
    StagedCells S;
    stage_step(G, Mo, dump, act, rec, ns - q, traj, stage_rows(sg_lds), S);
    if (!act) return;
    prep_step<F32, 1, StagedCells>(G, Mo, P, ns - q, rec, nt, dump, pkt + (size_t)g * stride, stride == PK,
                                   unit_scale(), ax, S);
  } else {
    if (!act) return;
    prep_step<F32>(G, Mo, P, ns - q, rec, nt, dump, pkt + (size_t)g * stride, stride == PK, unit_scale(), ax,
                   NoVG(), d32);
  }
}

/* One lane per ray: the recurrence of N and I over the ray's packets
 * (SEG = false: the rays with seg_base < 0) or the composition of its
 * segment maps (SEG = true: the rays with seg_base >= 0).  Two instances,
 * so that neither carries the other's registers. */
template <bool SEG>
__global__ __launch_bounds__(WAVE) void k_chain(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P, const int *__restrict__ pix, int nrays,
                                                int cap, int stride, const long long *__restrict__ off,
                                                const int *__restrict__ nstep_in,
                                                const double *__restrict__ traj,
                                                const double *__restrict__ dump,
                                                const double *__restrict__ pkt,
                                                const int *__restrict__ seg_base, int seg_len,
                                                const double *__restrict__ segmap,
                                                const int *__restrict__ segflag,
                                                double *__restrict__ out7, int *__restrict__ flags,
                                                int *__restrict__ status,
                                                unsigned long long *__restrict__ counts)
{
  int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= nrays) return;
  const int sb = seg_base ? seg_base[ray] : -1;
  if ((sb >= 0) != SEG) return;
  int pi, pj;
  if (!ray_pixel(pix, ray, P.nx, P.ny, pi, pj)) return;
  int ns = nstep_in[ray];
  double *o = out7 + (size_t)ray * 7;
  if (ns >= P.maxnstep - 1) {
    atomicMin(&status[1], pixel_code(pi, pj));
    for (int q = 0; q < 7; q++) o[q] = 0.;
    flags[ray] = 0;
    return;
  }
  const double *rec = traj + (size_t)ray * cap * REC;
  double js, ks;
  if constexpr (SEG) {
    if (P.transfer_fp32) far_end_coefficients<true>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
    else far_end_coefficients<false>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
    IPG_SCHED_BARRIER();
  } else {
    if (P.transfer_fp32) far_end_coefficients<true>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
    else far_end_coefficients<false>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
  }
  double res[7];
  int nrad;
  if constexpr (SEG)
    flags[ray] = chain_ray_segmented(G, P, ns, seg_len, segmap + (size_t)sb * SEGCOL * NPAR, segflag + sb, js, ks,
                                     res, nrad);
  else
    flags[ray] = chain_ray(G, P, ns, pkt + (size_t)off[ray] * stride, stride, js, ks, res, nrad);
  for (int q = 0; q < 7; q++) o[q] = res[q];
  atomicAdd(&counts[1], (unsigned long long)nrad);
}

/* save_pixel (src/main.c:885-910): reference layout, nu^3 scaling, QU sign */
__global__ void k_finalize(ipgpu_image P, const int *__restrict__ pix, int out_ny, int nrays, double f3,
                           const double *__restrict__ out7,
                           const int *__restrict__ flags_ray, const int *__restrict__ nstep_ray,
                           double *__restrict__ image, double *__restrict__ imageS,
                           double *__restrict__ taus, int *__restrict__ nstep_pix,
                           int *__restrict__ flag_pix, int *__restrict__ status)
{
  int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= nrays) return;
  int i, j;
  if (!ray_pixel(pix, ray, P.nx, P.ny, i, j)) return;
  size_t p = (size_t)i * out_ny + j;
  const double *o = out7 + (size_t)ray * 7;
  image[p] = o[0] * f3;
  taus[p] = o[5];
  nstep_pix[p] = nstep_ray[ray];
  flag_pix[p] = flags_ray[ray];
  if (!P.only_unpolarized) {
    double qs = (P.qu_conv == 0) ? -1. : 1.;
    imageS[p * 5 + 0] = o[1] * f3;
    imageS[p * 5 + 1] = qs * o[2] * f3;
    imageS[p * 5 + 2] = qs * o[3] * f3;
    imageS[p * 5 + 3] = o[4] * f3;
    imageS[p * 5 + 4] = o[6];
    if (isnan(imageS[p * 5])) atomicMin(&status[2], pixel_code(i, j));
  }
}

/* ------------------------------------------------------------ dump kernels */

__device__ __forceinline__ size_t gcell(const ipgpu_geom &G, int i, int j, int k)
{
  return ((size_t)i * (G.N2 + 2) + j) * (size_t)(G.N3 + 2) + k;
}

/* ghosted SoA fields -> 16-double cell records */
__global__ void k_pack(size_t ncell, const double *__restrict__ soa, double *__restrict__ dump)
{
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double *d = dump + CELL_STRIDE * c;
  for (int f = 0; f < IPGPU_NFIELDS; f++) d[f] = soa[f * ncell + c];
  for (int f = IPGPU_NFIELDS; f < CELL_STRIDE; f++) d[f] = 0.;
}

/* slots of the derivation: 10 rho, 11 u, 12 beta */
constexpr int S_RHO = 10, S_UU = 11, S_BETA = 12;

/* interior cells: primitives in, b from the local metric
 * (model/iharm/model.c:1954-2010) */
__global__ void k_derive_b(ipgpu_geom G, double B_unit, int nprim, const double *__restrict__ prims,
                           double *__restrict__ dump)
{
  size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t ntot = (size_t)G.N1 * G.N2 * G.N3;
  if (n >= ntot) return;
  int k = (int)(n % G.N3);
  int j = (int)((n / G.N3) % G.N2);
  int i = (int)(n / ((size_t)G.N3 * G.N2));
  const double *p = prims + n * nprim;
  double *d = dump + CELL_STRIDE * gcell(G, i + 1, j + 1, k + 1);
  Metric m;
  double X1 = G.startx[1] + (i + 0.5) * G.dx[1];
  double X2 = G.startx[2] + (j + 0.5) * G.dx[2];
  metric(G, X1, X2, m);
  double U[3] = {p[2], p[3], p[4]}, Bp[3] = {p[5], p[6], p[7]};
  double UdotU = 0.;
  for (int l = 1; l < 4; l++)
    for (int mm = 1; mm < 4; mm++) UdotU += m.g[sy(l, mm)] * U[l - 1] * U[mm - 1];
  double ufac = sqrt(-1. / m.gc[0] * (1 + fabs(UdotU)));
  double ucon[4], ucov[4], bcon[4], bcov[4];
  ucon[0] = -ufac * m.gc[0];
  for (int l = 1; l < 4; l++) ucon[l] = U[l - 1] - ufac * m.gc[sy(0, l)];
  lower(m.g, ucon, ucov);
  double udotB = 0.;
  for (int l = 1; l < 4; l++) udotB += ucov[l] * Bp[l - 1];
  bcon[0] = udotB;
  for (int l = 1; l < 4; l++) bcon[l] = (Bp[l - 1] + ucon[l] * udotB) / ucon[0];
  lower(m.g, bcon, bcov);
  double bsq = dot4(bcon, bcov);
  for (int f = 0; f < 6; f++) d[f] = p[2 + f];
  d[8] = sqrt(bsq) * B_unit;
  d[S_RHO] = p[0];
  d[S_UU] = p[1];
}

__device__ __forceinline__ void copy_ghost(double *dst, const double *src)
{
  for (int f = 0; f < 6; f++) dst[f] = src[f];
  dst[8] = src[8];
  dst[S_RHO] = src[S_RHO];
  dst[S_UU] = src[S_UU];
}

/* ghost zones, model/iharm/model.c:1314-1370: radial copy */
__global__ void k_bc_radial(ipgpu_geom G, double *dump)
{
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= G.N2 * G.N3) return;
  int j = n / G.N3 + 1, k = n % G.N3 + 1;
  copy_ghost(dump + CELL_STRIDE * gcell(G, 0, j, k), dump + CELL_STRIDE * gcell(G, 1, j, k));
  copy_ghost(dump + CELL_STRIDE * gcell(G, G.N1 + 1, j, k), dump + CELL_STRIDE * gcell(G, G.N1, j, k));
}

/* polar: rotation by pi (kflip) */
__global__ void k_bc_polar(ipgpu_geom G, double *dump)
{
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (G.N1 + 2) * G.N3) return;
  int i = n / G.N3, k = n % G.N3 + 1;
  const int N2 = G.N2, N3 = G.N3;
  if (N3 % 2 == 0) {
    int kflip = ((k - 1) + (N3 / 2)) % N3 + 1;
    copy_ghost(dump + CELL_STRIDE * gcell(G, i, 0, k), dump + CELL_STRIDE * gcell(G, i, 1, kflip));
    copy_ghost(dump + CELL_STRIDE * gcell(G, i, N2 + 1, k), dump + CELL_STRIDE * gcell(G, i, N2, kflip));
  } else {
    int k1 = (k + (N3 / 2)) % N3, k2 = (k + (N3 / 2) + 1) % N3;
    const double *a0 = dump + CELL_STRIDE * gcell(G, i, 1, k1), *a1 = dump + CELL_STRIDE * gcell(G, i, 1, k2);
    const double *b0 = dump + CELL_STRIDE * gcell(G, i, N2, k1), *b1 = dump + CELL_STRIDE * gcell(G, i, N2, k2);
    double *lo = dump + CELL_STRIDE * gcell(G, i, 0, k), *hi = dump + CELL_STRIDE * gcell(G, i, N2 + 1, k);
    const int fs[9] = {0, 1, 2, 3, 4, 5, 8, S_RHO, S_UU};
    for (int q = 0; q < 9; q++) {
      lo[fs[q]] = (a0[fs[q]] + a1[fs[q]]) / 2.;
      hi[fs[q]] = (b0[fs[q]] + b1[fs[q]]) / 2.;
    }
  }
}

/* azimuth: periodic */
__global__ void k_bc_phi(ipgpu_geom G, double *dump)
{
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (G.N1 + 2) * (G.N2 + 2)) return;
  int i = n / (G.N2 + 2), j = n % (G.N2 + 2);
  copy_ghost(dump + CELL_STRIDE * gcell(G, i, j, 0), dump + CELL_STRIDE * gcell(G, i, j, G.N3));
  copy_ghost(dump + CELL_STRIDE * gcell(G, i, j, G.N3 + 1), dump + CELL_STRIDE * gcell(G, i, j, 1));
}

/* n_e, Theta_e (R_high/R_low, ELECTRONS = 2), sigma, beta over all cells:
 * model/iharm/model.c:610-717 */
__global__ void k_derive_thermo(size_t ncell, ipgpu_units U, double RHO_unit, double B_unit,
                                double *__restrict__ dump)
{
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double *d = dump + CELL_STRIDE * c;
  double rho = d[S_RHO], uu = d[S_UU];
  d[6] = rho * RHO_unit / (MP + ME) * 1.;
  double bsq = d[8] / B_unit;
  bsq = bsq * bsq;
  double sigma_m = bsq / rho;
  double beta_m = uu * (U.gam - 1.) / 0.5 / bsq;
  double betasq = beta_m * beta_m / U.beta_crit / U.beta_crit;
  double trat = U.trat_large * betasq / (1. + betasq) + U.trat_small / (1. + betasq);
  double Tu = (MP / ME) * (U.game - 1.) * (U.gamp - 1.) / ((U.gamp - 1.) + (U.game - 1.) * trat);
  double th = Tu * uu / rho;
  d[7] = fmax(th, 1.e-3);
  d[9] = fmax(sigma_m, SMALL);
  d[S_BETA] = fmax(beta_m, SMALL);
  /* dTheta_e/dR_high and dR_low of this electron model (zero where the
   * 1e-3 floor holds), for gradients of images with respect to them */
  double dth_dtrat = (th > 1.e-3) ? -th * (U.game - 1.) / ((U.gamp - 1.) + (U.game - 1.) * trat) : 0.;
  d[SLOT_DTH_RHIGH] = dth_dtrat * betasq / (1. + betasq);
  d[SLOT_DTH_RLOW] = dth_dtrat / (1. + betasq);
}

__global__ void k_unpack(size_t ncell, const double *__restrict__ dump, double *__restrict__ soa11)
{
  size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const double *d = dump + CELL_STRIDE * c;
  for (int f = 0; f < 10; f++) soa11[f * ncell + c] = d[f];
  soa11[10 * ncell + c] = d[S_BETA];
}

/* ------------------------------------------------------------- batches */

/* Step-count bands of the batch transfer's item order (2^ITEM_BAND steps). */
constexpr int ITEM_BAND = 8;

/* A trace group: one camera/image geometry traced once for all its jobs. */
struct BGroup {
  ipgpu_image P;      /* image parameters at the trace frequency */
  long long ray_base; /* global index of the group's first ray */
  int nrays;          /* tiles * 64 */
  int pad;
};

/* A job of a batch on the device. */
struct BJob {
  ipgpu_image P;      /* the job's own image parameters */
  Scale sc;           /* relative to its group's trace and the uploaded dump */
  long long pix_off;  /* first pixel of the job in the batch output arrays */
  int group;
  int pad;
};

__device__ __forceinline__ int group_of(const BGroup *__restrict__ groups, int ngroups, long long r)
{
  int lo = 0, hi = ngroups;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (groups[mid].ray_base <= r) lo = mid;
    else hi = mid;
  }
  return lo;
}

/* Geodesic step counts of all rays of all groups (nothing stored). */
__global__ __launch_bounds__(WAVE) void k_trace_count(ipgpu_geom G, const BGroup *__restrict__ groups,
                                                      int ngroups, long long nr, int *__restrict__ nstep_out)
{
  long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const int g = group_of(groups, ngroups, r);
  const ipgpu_image P = groups[g].P;
  int pi, pj;
  ray_pixel(nullptr, (int)(r - groups[g].ray_base), P.nx, P.ny, pi, pj);
  if (pi >= P.nx || pj >= P.ny) {
    nstep_out[r] = 0;
    return;
  }
  nstep_out[r] = trace_ray(G, P, pi, pj, 0, nullptr, nullptr);
}

/* The same geodesics again, longest first, each written to its exact slice
 * of the compact record buffer (off: first record of the ray). */
__global__ __launch_bounds__(WAVE) void k_trace_store(ipgpu_geom G, const BGroup *__restrict__ groups,
                                                      int ngroups, const int *__restrict__ order, int nr,
                                                      const long long *__restrict__ off,
                                                      const int *__restrict__ nstep_in, double *__restrict__ traj,
                                                      int *__restrict__ nturns_buf, int *__restrict__ status,
                                                      unsigned long long *__restrict__ counts)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nr) return;
  const int r = order[t];
  const int g = group_of(groups, ngroups, r);
  const ipgpu_image P = groups[g].P;
  int pi, pj;
  ray_pixel(nullptr, (int)(r - groups[g].ray_base), P.nx, P.ny, pi, pj);
  const int ns = nstep_in[r];
  double *rec = traj + (size_t)off[r] * REC;
  int *nt = (nturns_buf && P.target_nturns >= 0) ? nturns_buf + off[r] : nullptr;
  int nstep = trace_ray(G, P, pi, pj, ns + 1, rec, nt);
  if (nstep != ns) atomicAdd(&status[0], 1);   /* the two passes are the same arithmetic */
  atomicAdd(&counts[0], (unsigned long long)nstep);
}

/* The batch's trace passes without subring counts run on lane pairs (two
 * lanes per ray, trace_pair), whose records a resumed trace continues bit
 * for bit.  Pass 1: every ray's step count and, with slot > 0, its first
 * slot records in a slot of its own (traj + r slot REC); a ray of fewer
 * than slot steps is then complete. */
__global__ __launch_bounds__(WAVE) void k_trace_slot(ipgpu_geom G, const BGroup *__restrict__ groups,
                                                     int ngroups, long long nr, int slot, double *__restrict__ traj,
                                                     int *__restrict__ nstep_out, const int *__restrict__ wtile,
                                                     double *__restrict__ arena, unsigned long long *ovf_counter,
                                                     unsigned long long nchunks, int *__restrict__ ovf_table,
                                                     int maxc, int *__restrict__ ovf_ok)
{
  /* a wave traces half a tile of 64 rays; wtile: the tiles in dispatch
   * order (null: in index order) */
  const int w = blockIdx.x, lane = threadIdx.x;
  const long long r = wtile ? (long long)wtile[w >> 1] * WAVE + (w & 1) * (WAVE / 2) + (lane >> 1)
                            : ((long long)w * WAVE + lane) >> 1;
  const bool odd = lane & 1;
  if (r >= nr) return;
  const int g = group_of(groups, ngroups, r);
  const ipgpu_image P = groups[g].P;
  int pi, pj;
  ray_pixel(nullptr, (int)(r - groups[g].ray_base), P.nx, P.ny, pi, pj);
  if (pi >= P.nx || pj >= P.ny) {
    if (!odd) nstep_out[r] = 0;
    return;
  }
  int ns;
  if (arena && slot > 0) {
    /* the records past the slot go to chunks of the arena as the ray runs */
    OverflowChunks ov{arena, ovf_counter, nchunks, ovf_table + (size_t)r * maxc, maxc, slot, nullptr, true};
    ns = trace_pair_slot(G, P, pi, pj, slot, traj + (size_t)r * slot * REC, odd, ov);
    if (!odd) ovf_ok[r] = ov.ok ? 1 : 0;
  } else {
    ns = trace_pair(G, P, pi, pj, slot, slot > 0 ? traj + (size_t)r * slot * REC : nullptr, 0, 0x7fffffff, nullptr,
                    odd);
  }
  if (!odd) nstep_out[r] = ns;
}

/* Pass 2 on lane pairs: the records of the nr rays of order in their
 * regions (offsets off, in records). */
__global__ __launch_bounds__(WAVE) void k_trace_store_pair(ipgpu_geom G, const BGroup *__restrict__ groups,
                                                           int ngroups, const int *__restrict__ order, int nr,
                                                           const long long *__restrict__ off,
                                                           const int *__restrict__ nstep_in,
                                                           double *__restrict__ traj, int *__restrict__ status,
                                                           unsigned long long *__restrict__ counts)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 1;
  const bool odd = t & 1;
  if (k >= nr) return;
  const int r = order[k];
  const int g = group_of(groups, ngroups, r);
  const ipgpu_image P = groups[g].P;
  int pi, pj;
  ray_pixel(nullptr, (int)(r - groups[g].ray_base), P.nx, P.ny, pi, pj);
  const int ns = nstep_in[r];
  const int nstep = trace_pair(G, P, pi, pj, ns + 1, traj + (size_t)off[r] * REC, 0, 0x7fffffff, nullptr, odd);
  if (odd) return;
  if (nstep != ns) atomicAdd(&status[0], 1);   /* the two passes are the same arithmetic */
  atomicAdd(&counts[0], (unsigned long long)nstep);
}

/* The slot records of the n rays of list (slot REC doubles each) copied to
 * their full-length regions in traj2 (record offsets off2[k]). */
__global__ void k_slot_copy(const int *__restrict__ list, int n, const long long *__restrict__ off2, int slot,
                            const double *__restrict__ traj, double *__restrict__ traj2)
{
  const long long per = (long long)slot * REC;
  const long long total = (long long)n * per;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(t / per);
    const long long e = t - (long long)k * per;
    traj2[off2[k] * REC + e] = traj[(long long)list[k] * per + e];
  }
}

/* The n rays of list whose records past the slot are all in arena chunks
 * (table: maxc chunk indices per ray): slot records and chunks copied to
 * their full-length regions (offsets off2, in records) of traj2, one block
 * per ray, 16 bytes per thread and step. */
__global__ void k_ovf_gather(const int *__restrict__ list, int n, const long long *__restrict__ off2, int slot,
                             const int *__restrict__ nstep, const double *__restrict__ traj,
                             const double *__restrict__ arena, const int *__restrict__ table, int maxc,
                             double *__restrict__ traj2)
{
  const int k = blockIdx.x;
  if (k >= n) return;
  const int r = list[k];
  const long long nq = (long long)(nstep[r] + 1) * (REC / 2);
  double2 *dst = reinterpret_cast<double2 *>(traj2 + (size_t)off2[k] * REC);
  for (long long u = threadIdx.x; u < nq; u += blockDim.x) {
    const int rn = (int)(u / (REC / 2)), part = (int)(u % (REC / 2));
    const double *src;
    if (rn < slot) {
      src = traj + ((size_t)r * slot + rn) * REC;
    } else {
      const int rel = rn - slot;
      src = arena + ((size_t)table[(size_t)r * maxc + rel / OVF_CH] * OVF_CH + (rel % OVF_CH)) * REC;
    }
    dst[u] = reinterpret_cast<const double2 *>(src)[part];
  }
}

/* The rays of list longer than the slot, on lane pairs: resumed from their
 * last slot record (slot - 1) in their full-length regions of traj2; a step
 * count other than the first pass's is reported in status[0]. */
__global__ __launch_bounds__(WAVE) void k_trace_resume_list(ipgpu_geom G, const BGroup *__restrict__ groups,
                                                            int ngroups, const int *__restrict__ list, int n,
                                                            const long long *__restrict__ off2, int slot,
                                                            const int *__restrict__ nstep_in,
                                                            double *__restrict__ traj2, int *__restrict__ status)
{
  IPG_LATENCY_CRITICAL();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 1;
  const bool odd = t & 1;
  if (k >= n) return;
  const int r = list[k];
  const int g = group_of(groups, ngroups, r);
  const ipgpu_image P = groups[g].P;
  const int ns = nstep_in[r];
  const int nstep = trace_pair(G, P, 0, 0, ns + 1, traj2 + (size_t)off2[k] * REC, slot - 1, 0x7fffffff, nullptr, odd);
  if (!odd && nstep != ns) atomicAdd(&status[0], 1);
}

/* The transfer of work item w (job, ray) and the job's pixel in the
 * reference layout; mi: the lane's LDS slot for the carried transport
 * matrix (FP64) or null. */
template <bool F32>
__device__ __forceinline__ void xfer_item(const ipgpu_geom &G, const ipgpu_model &Mo, const BJob *__restrict__ jobs,
                                          const BGroup *__restrict__ groups, const int2 *__restrict__ items, int w,
                                          const long long *__restrict__ off, const double *__restrict__ traj,
                                          const double *__restrict__ traj2, const int *__restrict__ nstep_in,
                                          const int *__restrict__ nturns_buf, const double *__restrict__ dump,
                                          double *__restrict__ image, double *__restrict__ imageS,
                                          double *__restrict__ taus, int *__restrict__ nstep_pix,
                                          int *__restrict__ flag_pix, unsigned long long *__restrict__ bstatus,
                                          unsigned long long *__restrict__ counts, double *mi,
                                          const float *__restrict__ d32)
{
  const int2 it = items[w];
  const int jb = it.x, r = it.y;
  const BJob J = jobs[jb];
  const ipgpu_image &P = J.P;
  int pi, pj;
  ray_pixel(nullptr, (int)(r - groups[J.group].ray_base), P.nx, P.ny, pi, pj);
  const long long pix = (long long)pi * P.ny + pj;
  const size_t p = (size_t)(J.pix_off + pix);
  const int ns = nstep_in[r];
  nstep_pix[p] = ns;
  if (ns >= P.maxnstep - 1) {
    atomicMin(&bstatus[0], ((unsigned long long)jb << 32 | (unsigned long long)pix) + 1ull);
    image[p] = 0.;
    taus[p] = 0.;
    flag_pix[p] = 0;
    if (!P.only_unpolarized)
      for (int q = 0; q < 5; q++) imageS[p * 5 + q] = 0.;
    return;
  }
  /* off[r] >= 0: records at traj + off[r] REC; < 0: at traj2 + (-1 - off[r]) REC */
  const long long ro = off[r];
  const double *rec = ro >= 0 ? traj + (size_t)ro * REC : traj2 + (size_t)(-1 - ro) * REC;
  const int *nt = (nturns_buf && P.target_nturns >= 0) ? nturns_buf + ro : nullptr;
  double o[7];
  int nrad;
  int flag = transfer_fused<F32>(G, Mo, P, ns, rec, nt, dump, J.sc, o, nrad, mi, d32);
  const double f3 = P.freqcgs * P.freqcgs * P.freqcgs;
  image[p] = o[0] * f3;
  taus[p] = o[5];
  flag_pix[p] = flag;
  if (!P.only_unpolarized) {
    double qs = (P.qu_conv == 0) ? -1. : 1.;
    imageS[p * 5 + 0] = o[1] * f3;
    imageS[p * 5 + 1] = qs * o[2] * f3;
    imageS[p * 5 + 2] = qs * o[3] * f3;
    imageS[p * 5 + 3] = o[4] * f3;
    imageS[p * 5 + 4] = o[6];
    if (isnan(o[1])) atomicMin(&bstatus[1], ((unsigned long long)jb << 32 | (unsigned long long)pix) + 1ull);
  }
  atomicAdd(&counts[1], (unsigned long long)nrad);
}

/* One lane per (job, ray) work item: the fused transfer of the ray for the
 * job, then the job's pixel in the reference layout (save_pixel,
 * src/main.c:885-910).  bstatus[0]: min of (job << 32 | pixel) + 1 over
 * rays past maxnstep; bstatus[1]: the same over NaN pixels. */
template <bool F32>
__global__ __launch_bounds__(WAVE) void k_xfer_batch(ipgpu_geom G, ipgpu_model Mo, const BJob *__restrict__ jobs,
                                                     const BGroup *__restrict__ groups,
                                                     const int2 *__restrict__ items, int nitems,
                                                     const long long *__restrict__ off,
                                                     const double *__restrict__ traj,
                                                     const double *__restrict__ traj2,
                                                     const int *__restrict__ nstep_in,
                                                     const int *__restrict__ nturns_buf,
                                                     const double *__restrict__ dump, double *__restrict__ image,
                                                     double *__restrict__ imageS, double *__restrict__ taus,
                                                     int *__restrict__ nstep_pix, int *__restrict__ flag_pix,
                                                     unsigned long long *__restrict__ bstatus,
                                                     unsigned long long *__restrict__ counts,
                                                     const float *__restrict__ d32)
{
  /* FP64: the transport matrix carried between steps lives in LDS (one
   * padded slot per lane), which keeps the step within the register file */
  __shared__ double mi_lds[F32 ? 1 : WAVE * 17];
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nitems) return;
  xfer_item<F32>(G, Mo, jobs, groups, items, w, off, traj, traj2, nstep_in, nturns_buf, dump, image, imageS, taus,
                 nstep_pix, flag_pix, bstatus, counts, F32 ? nullptr : mi_lds + 17 * threadIdx.x, d32);
}

/* The same work as a persistent launch with a work queue: a fixed number of
 * waves, each taking the next chunk of 64 items (an atomic ticket on
 * queue[0]) until the list is done.  Items are ordered longest ray first,
 * so the waves that finish early take the remaining short items. */
template <bool F32>
__global__ __launch_bounds__(WAVE) void k_xfer_queue(ipgpu_geom G, ipgpu_model Mo, const BJob *__restrict__ jobs,
                                                     const BGroup *__restrict__ groups,
                                                     const int2 *__restrict__ items, int nitems,
                                                     unsigned int *__restrict__ queue,
                                                     const long long *__restrict__ off,
                                                     const double *__restrict__ traj,
                                                     const double *__restrict__ traj2,
                                                     const int *__restrict__ nstep_in,
                                                     const int *__restrict__ nturns_buf,
                                                     const double *__restrict__ dump, double *__restrict__ image,
                                                     double *__restrict__ imageS, double *__restrict__ taus,
                                                     int *__restrict__ nstep_pix, int *__restrict__ flag_pix,
                                                     unsigned long long *__restrict__ bstatus,
                                                     unsigned long long *__restrict__ counts,
                                                     const float *__restrict__ d32)
{
  __shared__ double mi_lds[F32 ? 1 : WAVE * 17];
  __shared__ unsigned int ticket;
  const unsigned int nchunks = (unsigned int)((nitems + WAVE - 1) / WAVE);
  while (true) {
    if (threadIdx.x == 0) ticket = atomicAdd(&queue[0], 1u);
    __syncthreads();
    const unsigned int c = ticket;
    __syncthreads();
    if (c >= nchunks) break;   /* every lane of the wave leaves together */
    const int w = (int)(c * WAVE + threadIdx.x);
    if (w < nitems)
      xfer_item<F32>(G, Mo, jobs, groups, items, w, off, traj, traj2, nstep_in, nturns_buf, dump, image, imageS,
                     taus, nstep_pix, flag_pix, bstatus, counts, F32 ? nullptr : mi_lds + 17 * threadIdx.x, d32);
  }
}

/* The transfer of one ray for a subset of the jobs of its trace group
 * (transfer_fused_multi): item (ray, subset), subset {n, job0, job1, job2}. */
template <bool F32>
__device__ __forceinline__ void xfer_multi_item(const ipgpu_geom &G, const ipgpu_model &Mo, const BJob *jobs,
                                                const BGroup *groups, const int2 *items, const int4 *subsets,
                                                int w, const long long *off, const double *traj,
                                                const double *traj2, const int *nstep_in, const int *nturns_buf,
                                                const double *dump, double *image, double *imageS, double *taus,
                                                int *nstep_pix, int *flag_pix, unsigned long long *bstatus,
                                                unsigned long long *counts, unsigned long long *bcounts, double *mi,
                                                double *nst, const float *d32)
{
  const int2 it = items[w];
  const int r = it.x;
  const int4 sb = subsets[it.y];
  const int nj = sb.x;
  const int jbs[MJ] = {sb.y, sb.z, sb.w};
  const ipgpu_image &P0 = jobs[jbs[0]].P;
  int pi, pj;
  ray_pixel(nullptr, (int)(r - groups[jobs[jbs[0]].group].ray_base), P0.nx, P0.ny, pi, pj);
  const long long pix = (long long)pi * P0.ny + pj;
  const int ns = nstep_in[r];
  if (ns >= P0.maxnstep - 1) {
#pragma unroll
    for (int j = 0; j < MJ; j++) {
      if (j >= nj) break;
      const BJob &J = jobs[jbs[j]];
      const size_t p = (size_t)(J.pix_off + pix);
      atomicMin(&bstatus[0], ((unsigned long long)jbs[j] << 32 | (unsigned long long)pix) + 1ull);
      nstep_pix[p] = ns;
      image[p] = 0.;
      taus[p] = 0.;
      flag_pix[p] = 0;
      if (!J.P.only_unpolarized)
        for (int q = 0; q < 5; q++) imageS[p * 5 + q] = 0.;
    }
    return;
  }
  /* off[r] >= 0: records at traj + off[r] REC; < 0: at traj2 + (-1 - off[r]) REC */
  const long long ro = off[r];
  const double *rec = ro >= 0 ? traj + (size_t)ro * REC : traj2 + (size_t)(-1 - ro) * REC;
  const int *nt = (nturns_buf && P0.target_nturns >= 0) ? nturns_buf + ro : nullptr;
  JobCfg cfg[MJ];
  const ipgpu_image *Pp[MJ];
#pragma unroll
  for (int j = 0; j < MJ; j++) {
    const BJob &J = jobs[jbs[j < nj ? j : 0]];
    cfg[j].sc = J.sc;
    cfg[j].etype = J.P.emission_type;
    cfg[j].pol = !J.P.only_unpolarized;
    cfg[j].floors = J.P.stokes_floors;
    cfg[j].cj = J.P.counterjet;
    Pp[j] = &J.P;
  }
  double o[7 * MJ];
  int fl[MJ], nr[MJ];
  transfer_fused_multi<F32>(G, Mo, nj, cfg, Pp, ns, rec, nt, dump, o, fl, nr, mi, nst, WAVE, d32);
  unsigned long long nrad = 0;
#pragma unroll
  for (int j = 0; j < MJ; j++) {
    if (j >= nj) break;
    const BJob &J = jobs[jbs[j]];
    const ipgpu_image &P = J.P;
    const size_t p = (size_t)(J.pix_off + pix);
    const double f3 = P.freqcgs * P.freqcgs * P.freqcgs;
    const double *oj = o + 7 * j;
    nstep_pix[p] = ns;
    image[p] = oj[0] * f3;
    taus[p] = oj[5];
    flag_pix[p] = fl[j];
    if (!P.only_unpolarized) {
      const double qs = (P.qu_conv == 0) ? -1. : 1.;
      imageS[p * 5 + 0] = oj[1] * f3;
      imageS[p * 5 + 1] = qs * oj[2] * f3;
      imageS[p * 5 + 2] = qs * oj[3] * f3;
      imageS[p * 5 + 3] = oj[4] * f3;
      imageS[p * 5 + 4] = oj[6];
      if (isnan(oj[1])) atomicMin(&bstatus[1], ((unsigned long long)jbs[j] << 32 | (unsigned long long)pix) + 1ull);
    }
    nrad += (unsigned long long)nr[j];
  }
  atomicAdd(&counts[1], nrad);
  atomicAdd(&bcounts[0], (unsigned long long)nr[0]);   /* the ray's radiating steps, once */
}

/* One lane per (ray, job subset) item, in waves of 64 items or, with a
 * queue, as a persistent launch taking chunks of 64 items by an atomic
 * ticket.  The transport matrix carried between steps and the jobs'
 * coherency tensors live in LDS (a lane's slots, lane-minor). */
template <bool F32>
__global__ __launch_bounds__(WAVE) void k_xfer_multi(ipgpu_geom G, ipgpu_model Mo, const BJob *__restrict__ jobs,
                                                     const BGroup *__restrict__ groups,
                                                     const int2 *__restrict__ items, const int4 *__restrict__ subsets,
                                                     int nitems, unsigned int *__restrict__ queue,
                                                     const long long *__restrict__ off,
                                                     const double *__restrict__ traj,
                                                     const double *__restrict__ traj2,
                                                     const int *__restrict__ nstep_in,
                                                     const int *__restrict__ nturns_buf,
                                                     const double *__restrict__ dump, double *__restrict__ image,
                                                     double *__restrict__ imageS, double *__restrict__ taus,
                                                     int *__restrict__ nstep_pix, int *__restrict__ flag_pix,
                                                     unsigned long long *__restrict__ bstatus,
                                                     unsigned long long *__restrict__ counts,
                                                     unsigned long long *__restrict__ bcounts,
                                                     const float *__restrict__ d32)
{
  __shared__ double mi_lds[WAVE * 17];
  __shared__ double nst_lds[MJ * 16 * WAVE];
  __shared__ unsigned int ticket;
  double *mi = mi_lds + 17 * threadIdx.x;
  double *nst = nst_lds + threadIdx.x;
  if (!queue) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nitems) return;
    xfer_multi_item<F32>(G, Mo, jobs, groups, items, subsets, w, off, traj, traj2, nstep_in, nturns_buf, dump,
                         image, imageS, taus, nstep_pix, flag_pix, bstatus, counts, bcounts, mi, nst, d32);
    return;
  }
  const unsigned int nchunks = (unsigned int)((nitems + WAVE - 1) / WAVE);
  while (true) {
    if (threadIdx.x == 0) ticket = atomicAdd(&queue[0], 1u);
    __syncthreads();
    const unsigned int c = ticket;
    __syncthreads();
    if (c >= nchunks) break;   /* every lane of the wave leaves together */
    const int w = (int)(c * WAVE + threadIdx.x);
    if (w < nitems)
      xfer_multi_item<F32>(G, Mo, jobs, groups, items, subsets, w, off, traj, traj2, nstep_in, nturns_buf, dump,
                           image, imageS, taus, nstep_pix, flag_pix, bstatus, counts, bcounts, mi, nst, d32);
  }
}

/* The FP32 copy of the cell records the mixed-precision batch transfer
 * gathers: the 10 fields of each of the nc cells rounded to nearest, the
 * other 6 floats of its 64 bytes zero. */
__global__ void k_dump_f32(long long nc, const double *__restrict__ d, float *__restrict__ o)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  for (int f = 0; f < CELL32; f++) o[i * CELL32 + f] = f < 10 ? (float)d[i * CELL_STRIDE + f] : 0.f;
}

/* queue[0] = 0; with bstatus, both status words to "none"; with counts,
 * the radiating step count to 0 (a kernel rather than memsets, so that the
 * clears can sit in a captured graph) */
__global__ void k_queue_reset(unsigned int *queue, unsigned long long *bstatus, unsigned long long *counts,
                              unsigned long long *bcounts)
{
  if (threadIdx.x == 0) {
    queue[0] = 0u;
    if (bstatus) {
      bstatus[0] = ~0ull;
      bstatus[1] = ~0ull;
    }
    if (counts) counts[1] = 0ull;
    if (bcounts) bcounts[0] = 0ull;
  }
}

/* ------------------------------------------------------------- adjoint */

/* The reverse pass in segments (parallel in time).  Segment s of the
 * adjoint covers packets [seg_q0[s], seg_q1[s]) of ray seg_ray[s]; ray r's
 * segments are ray_seg0[r] .. ray_seg0[r+1] - 1, in order from the far end.
 *
 * One lane per (segment, column) of the segments' maps: column lane_col[l]
 * of segment lane_seg[l] (0..15 from the basis states, 16 the source term;
 * a ray's first segment has its source column only). */
__global__ __launch_bounds__(WAVE) void k_adj_map(int nlanes, const int *__restrict__ lane_seg,
                                                  const int *__restrict__ lane_col, const int *__restrict__ seg_ray,
                                                  const int *__restrict__ seg_q0, const int *__restrict__ seg_q1,
                                                  const long long *__restrict__ off, const double *__restrict__ pkt,
                                                  double *__restrict__ map)
{
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlanes) return;
  const int sg = lane_seg[l], b = lane_col[l];
  const int r = seg_ray[sg];
  Herm N;
  segment_column<1>(pkt + (size_t)off[r] * PK, seg_q0[sg], seg_q1[sg], b, N);
  double *o = map + (size_t)sg * ASEG_MAP + (size_t)b * NPAR;
#pragma unroll
  for (int p = 0; p < NPAR; p++) o[p] = herm_par(N, p);
}

/* One lane per ray: the composition of its segments' maps (adjoint_compose):
 * states before its segments, its pixel's I, Q, U, V and loss, adjoint
 * states after its segments.  cot (imageS layout) replaces the L2 loss
 * against target with weights w when not null. */
__global__ __launch_bounds__(WAVE) void k_adj_compose(ipgpu_geom G, ipgpu_image P, int nrays,
                                                      const int *__restrict__ nstep_in,
                                                      const int *__restrict__ ray_seg0,
                                                      const double *__restrict__ map,
                                                      const double *__restrict__ target, AdjW w,
                                                      const double *__restrict__ cot, Herm *__restrict__ Nstart,
                                                      Herm *__restrict__ Lend, double *__restrict__ loss,
                                                      double *__restrict__ imageS, int *__restrict__ status)
{
  IPG_LATENCY_CRITICAL();
  const int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= nrays) return;
  loss[ray] = 0.;
  int pi, pj;
  if (!ray_pixel(nullptr, ray, P.nx, P.ny, pi, pj)) return;
  const int ns = nstep_in[ray];
  const size_t p = (size_t)pi * P.ny + pj;
  if (ns >= P.maxnstep - 1) {
    atomicMin(&status[1], (int)p + 1);
    return;
  }
  double tg[4] = {0., 0., 0., 0.}, cv[4] = {0., 0., 0., 0.};
  for (int q = 0; q < 4; q++) {
    if (cot) cv[q] = cot[5 * p + q];
    else tg[q] = target[5 * p + q];
  }
  const int s0 = ray_seg0[ray], nseg = ray_seg0[ray + 1] - s0;
  double S[4];
  loss[ray] = adjoint_compose(G, P, nseg, map + (size_t)s0 * ASEG_MAP, tg, w.w, cot ? cv : nullptr, Nstart + s0,
                              Lend + s0, S);
  for (int q = 0; q < 4; q++) imageS[5 * p + q] = S[q];
  imageS[5 * p + 4] = 0.;
}

/* One lane per segment: its reverse pass (adjoint_range) from the state
 * before it and the adjoint state after it; g3 per segment, per-cell
 * gradients by atomics. */
__global__ __launch_bounds__(WAVE) void k_adj_seg(ipgpu_geom G, ipgpu_image P, int nseg,
                                                  const int *__restrict__ seg_ray, const int *__restrict__ seg_q0,
                                                  const int *__restrict__ seg_q1, const long long *__restrict__ off,
                                                  const double *__restrict__ pkt, const double *__restrict__ aux,
                                                  const double *__restrict__ dump, const Herm *__restrict__ Nstart,
                                                  const Herm *__restrict__ Lend, int nck, Herm *__restrict__ ckpt,
                                                  Herm *__restrict__ segbuf, double *__restrict__ gcell,
                                                  double *__restrict__ g3)
{
  IPG_LATENCY_CRITICAL();
  const int sg = blockIdx.x * blockDim.x + threadIdx.x;
  if (sg >= nseg) return;
  const int r = seg_ray[sg];
  Herm L = Lend[sg];
  double g[3] = {0., 0., 0.};
  adjoint_range(G, P, pkt + (size_t)off[r] * PK, aux + (size_t)off[r] * AX, seg_q0[sg], seg_q1[sg], Nstart[sg], L,
                dump, ckpt + (size_t)sg * nck, segbuf + (size_t)sg * CK, gcell, g);
  for (int q = 0; q < 3; q++) g3[(size_t)sg * 3 + q] = g[q];
}

/* n_e of interior cell (i, j, k) times scale[i][j][k] */
__global__ void k_scale_ne(ipgpu_geom G, const double *__restrict__ scale, double *__restrict__ dump)
{
  size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (size_t)G.N1 * G.N2 * G.N3) return;
  int k = (int)(n % G.N3), j = (int)((n / G.N3) % G.N2), i = (int)(n / ((size_t)G.N3 * G.N2));
  dump[CELL_STRIDE * gcell(G, i + 1, j + 1, k + 1) + 6] *= scale[n];
}

/* the ghost-zone rules of the derivation for even N3, for n_e alone (n_e
 * is linear in rho, so a ghost's n_e follows its source's as rho does) */
__global__ void k_bc_ne(ipgpu_geom G, int stage, double *dump)
{
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int N1 = G.N1, N2 = G.N2, N3 = G.N3;
  auto NE = [&](int i, int j, int k) -> double & { return dump[CELL_STRIDE * gcell(G, i, j, k) + 6]; };
  if (stage == 0) {
    if (n >= N2 * N3) return;
    int j = n / N3 + 1, k = n % N3 + 1;
    NE(0, j, k) = NE(1, j, k);
    NE(N1 + 1, j, k) = NE(N1, j, k);
  } else if (stage == 1) {
    if (n >= (N1 + 2) * N3) return;
    int i = n / N3, k = n % N3 + 1;
    int kflip = ((k - 1) + (N3 / 2)) % N3 + 1;   /* even N3 (checked by the caller) */
    NE(i, 0, k) = NE(i, 1, kflip);
    NE(i, N2 + 1, k) = NE(i, N2, kflip);
  } else {
    if (n >= (N1 + 2) * (N2 + 2)) return;
    int i = n / (N2 + 2), j = n % (N2 + 2);
    NE(i, j, 0) = NE(i, j, N3);
    NE(i, j, N3 + 1) = NE(i, j, 1);
  }
}

/* ------------------------------------------------ streamed single image */

/* First trace phase: every ray with a step budget; live[ray] = 1 where it
 * paused. */
__global__ __launch_bounds__(WAVE) void k_trace_budget(ipgpu_geom G, ipgpu_image P, const int *__restrict__ pix, int nrays, int cap, int nend,
                                                       double *__restrict__ traj, int *__restrict__ nstep_out,
                                                       int *__restrict__ live, int *__restrict__ host_ns,
                                                       int *__restrict__ host_live,
                                                       unsigned long long *__restrict__ counts)
{
  IPG_LATENCY_CRITICAL();
  /* one lane pair per ray */
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ray = t >> 1;
  const bool odd = t & 1;
  if (ray >= nrays) return;
  int pi, pj;
  int nstep = 0, lv = 0;
  if (ray_pixel(pix, ray, P.nx, P.ny, pi, pj)) {
    int fin = 1;
    nstep = trace_pair(G, P, pi, pj, cap, traj + (size_t)ray * cap * REC, 0, nend, &fin, odd);
    lv = fin ? 0 : 1;
    if (!odd) atomicAdd(&counts[0], (unsigned long long)nstep);
  }
  if (odd) return;
  nstep_out[ray] = nstep;
  live[ray] = lv;
  host_ns[ray] = nstep;
  host_live[ray] = lv;
}

/* A later trace phase over all rays: the rays still live resume from their
 * last record for at most nend steps in all (nend < 0: to the end).  The
 * host rows receive the step count of every ray integrated here and the
 * live flag of every ray. */
__global__ __launch_bounds__(WAVE) void k_trace_resume_mask(ipgpu_geom G, ipgpu_image P, int nrays, int cap,
                                                            int nend, double *__restrict__ traj,
                                                            int *__restrict__ nstep_io, int *__restrict__ live,
                                                            int *__restrict__ host_ns, int *__restrict__ host_live,
                                                            unsigned long long *__restrict__ counts,
                                                            const long long *__restrict__ rbase,
                                                            double *__restrict__ trajO, int ocap)
{
  IPG_LATENCY_CRITICAL();
  /* one lane pair per ray */
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ray = t >> 1;
  const bool odd = t & 1;
  if (ray >= nrays) return;
  if (!live[ray]) {
    if (!odd) host_live[ray] = 0;   /* the host row is complete for every ray */
    return;
  }
  const int n0 = nstep_io[ray];
  int fin = 1;
  /* a ray with a pooled region (the unbounded phase) continues there, with
   * the region's capacity; the others in their slots */
  double *rec = traj + (size_t)ray * cap * REC;
  int rcap = cap;
  if (rbase) {
    const long long b = rbase[ray];
    if (b >= 0) {
      rec = trajO + b;
      rcap = ocap;
    }
  }
  int nstep = trace_pair(G, P, 0, 0, rcap, rec, n0, nend < 0 ? 0x7fffffff : nend, &fin, odd);
  if (odd) return;
  nstep_io[ray] = nstep;
  live[ray] = fin ? 0 : 1;
  host_ns[ray] = nstep;
  host_live[ray] = fin ? 0 : 1;
  atomicAdd(&counts[0], (unsigned long long)(nstep - n0));
}

/* When the last bounded phase has ended: each ray still live takes, in
 * order of arrival, one of nreg full-length regions of trajO (reg doubles
 * each) for its records, or none once they are used up.  rbase of every ray
 * is written here (region offset in doubles, or -1), and its host row;
 * list[s] is the ray of region s. */
__global__ void k_pool_assign(int nrays, const int *__restrict__ live, int nreg, long long reg,
                              unsigned int *__restrict__ ctr, long long *__restrict__ rbase,
                              long long *__restrict__ host_rbase, int *__restrict__ list)
{
  const int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= nrays) return;
  long long b = -1;
  if (live[ray]) {
    const unsigned int s = atomicAdd(ctr, 1u);
    if (s < (unsigned int)nreg) {
      b = (long long)s * reg;
      list[s] = ray;
    }
  }
  rbase[ray] = b;
  host_rbase[ray] = b;
}

/* One workgroup per assigned region: the ray's records so far (at most cap,
 * its slot) copied from its slot in traj to its region, where the unbounded
 * phase resumes it and writes the rest. */
__global__ __launch_bounds__(256) void k_pool_copy(int nreg, const unsigned int *__restrict__ ctr,
                                                   const int *__restrict__ list, const int *__restrict__ nstep,
                                                   int cap, const double *__restrict__ traj,
                                                   double *__restrict__ trajO, long long reg)
{
  const unsigned int used = min(*ctr, (unsigned int)nreg);
  const unsigned int s = blockIdx.x;
  if (s >= used) return;
  const int ray = list[s];
  const int nrec = min(nstep[ray] + 1, cap);
  const long long n = (long long)nrec * REC / 2;      /* 16-byte pairs */
  const double2 *src = reinterpret_cast<const double2 *>(traj + (size_t)ray * cap * REC);
  double2 *dst = reinterpret_cast<double2 *>(trajO + (size_t)s * reg);
  for (long long i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

/* The rays of a single image whose trace outgrew the record capacity (the
 * unbounded phase counts their steps but stores only cap records): their
 * first cap records copied to full-length regions of trajO (record offsets
 * off[k]) by k_slot_copy, then resumed here on lane pairs from record
 * cap - 1 to the end, bit for bit the continuous trace.  rbase[ray]
 * receives the region's offset in doubles; a step count other than the
 * trace's is reported in status[0]. */
__global__ __launch_bounds__(WAVE) void k_trace_continue(ipgpu_geom G, ipgpu_image P, const int *__restrict__ list,
                                                         int n, const long long *__restrict__ off, int cap,
                                                         const int *__restrict__ nstep_in, double *__restrict__ trajO,
                                                         long long *__restrict__ rbase, int *__restrict__ status)
{
  IPG_LATENCY_CRITICAL();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t >> 1;
  const bool odd = t & 1;
  if (k >= n) return;
  const int r = list[k];
  const int ns = nstep_in[r];
  const int nstep = trace_pair(G, P, 0, 0, ns + 1, trajO + (size_t)off[k] * REC, cap - 1, 0x7fffffff, nullptr, odd);
  if (odd) return;
  rbase[r] = off[k] * REC;
  if (nstep != ns) atomicAdd(&status[0], 1);
}

/* Packets of a list of rays, wave-interleaved: the list is sorted by step
 * count, longest first, and cut into waves of 64 elements; wave c owns
 * rows W[c] .. W[c+1] (its first element's step count), and field f of step
 * q of its lane l is at pkt[((W[c] + q) * stride + f) * 64 + l].  A wave of
 * the chain kernel then loads one field of its 64 rays' packets as 512
 * contiguous bytes, and a wave of the packet kernel stores it so. */
constexpr int ILV = WAVE;

/* The packet kernel of an interleaved list works in tiles of 8 rays x 8
 * consecutive steps (one hardware wave): a lane's neighbours are the next
 * steps of its ray (the same trajectory lines, mostly the same dump cells)
 * and a store of one field covers 8 rows of 64 contiguous bytes.  Wave c of
 * the list has 8 ceil(len_c / 8) tiles, from tile TT[c] on. */
constexpr int TQ = 8;                        /* steps per tile */
constexpr int TL = WAVE / TQ;                /* rays per tile */
constexpr int TPB = PREP_BLOCK / WAVE;       /* tiles per block */

/* Block (TPB tiles) -> wave of the list whose tiles contain the block's
 * first tile. */
__global__ void k_blockmap_tiles(int nw, const long long *__restrict__ TT, int *__restrict__ blk)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nw) return;
  long long b0 = (TT[c] + TPB - 1) / TPB, b1 = (TT[c + 1] + TPB - 1) / TPB;
  for (long long b = b0; b < b1; b++) blk[b] = c;
}

/* Step packets of an interleaved list: one lane per (step, ray) slot of a
 * tile; slots past a ray's step count are padding and do nothing.  STAGE:
 * the dump gathers staged in LDS per 8-lane group (stage_cells), for which
 * every lane of a wave stays until its group's loads are issued. */
template <bool F32, bool STAGE>
__global__ __launch_bounds__(PREP_BLOCK) void k_prep_list(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P,
                                                          const int *__restrict__ list, int n,
                                                          const long long *__restrict__ W,
                                                          const long long *__restrict__ TT, int nw, long long tiles,
                                                          const int *__restrict__ blk, int cap, int stride,
                                                          const int *__restrict__ nstep_in,
                                                          const double *__restrict__ traj,
                                                          const long long *__restrict__ rbase,
                                                          const double *__restrict__ trajO,
                                                          const double *__restrict__ dump, double *__restrict__ pkt,
                                                          const float *__restrict__ d32)
{
  __shared__ double sg_lds[STAGE ? TPB * SG_WAVE_DBL : 1];
  const long long t = (long long)blockIdx.x * PREP_BLOCK + threadIdx.x;
  const long long tile = t / WAVE;
  const int j = (int)(t % WAVE);
  if (tile >= tiles) return;   /* a whole wave */
  int c = blk[blockIdx.x];
  while (c + 1 < nw && TT[c + 1] <= tile) c++;
  const int lt = (int)(tile - TT[c]);
  const int l = (lt % (ILV / TL)) * TL + j / TQ;   /* lane of the list wave */
  const int q = (lt / (ILV / TL)) * TQ + j % TQ;   /* step */
  const int k = c * ILV + l;
  bool act = k < n;
  const int ray = act ? list[k] : 0;
  const int ns = act ? nstep_in[ray] : 0;
  act = act && q < ns && ns < P.maxnstep - 1;
  const double *rec = act ? ray_rec(traj, cap, rbase, trajO, ray) : nullptr;
  double *pk = pkt + (size_t)(W[c] + q) * stride * ILV + pk_lane<ILV>(l);
  if constexpr (STAGE) {
    StagedCells S;
    stage_step(G, Mo, dump, act, rec, ns - q, traj, stage_rows(sg_lds), S);
    if (!act) return;
    prep_step<F32, ILV, StagedCells>(G, Mo, P, ns - q, rec, nullptr, dump, pk, stride == PK, unit_scale(), nullptr, S);
  } else {
    if (!act) return;
    prep_step<F32, ILV>(G, Mo, P, ns - q, rec, nullptr, dump, pk, stride == PK, unit_scale(), nullptr, NoVG(), d32);
  }
}

/* The same packets in the compact layout (ray after ray, packets of one ray
 * contiguous: element k's packets at offL[k] .. offL[k+1]), one lane per
 * step; for lists whose chains are all composed from segments, whose
 * column lanes then read one ray's packets from few cache lines. */
__global__ void k_blockmap_list(int n, const long long *__restrict__ offL, int *__restrict__ blk)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  long long b0 = (offL[k] + PREP_BLOCK - 1) / PREP_BLOCK, b1 = (offL[k + 1] + PREP_BLOCK - 1) / PREP_BLOCK;
  for (long long b = b0; b < b1; b++) blk[b] = k;
}

template <bool F32, bool STAGE>
__global__ __launch_bounds__(PREP_BLOCK) void k_prep_compact(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P,
                                                             const int *__restrict__ list, int n,
                                                             const long long *__restrict__ offL, long long total,
                                                             const int *__restrict__ blk, int cap,
                                                             const int *__restrict__ nstep_in,
                                                             const double *__restrict__ traj,
                                                             const long long *__restrict__ rbase,
                                                             const double *__restrict__ trajO,
                                                             const double *__restrict__ dump, double *__restrict__ pkt,
                                                             const float *__restrict__ d32)
{
  __shared__ double sg_lds[STAGE ? TPB * SG_WAVE_DBL : 1];
  long long g = (long long)blockIdx.x * PREP_BLOCK + threadIdx.x;
  bool act = g < total;
  int k = blk[blockIdx.x];
  if (act)
    while (k + 1 < n && offL[k + 1] <= g) k++;
  const int ray = list[k];
  const int ns = nstep_in[ray];
  act = act && ns < P.maxnstep - 1;
  const int q = (int)(g - offL[k]);
  const double *rec = act ? ray_rec(traj, cap, rbase, trajO, ray) : nullptr;
  if constexpr (STAGE) {
    StagedCells S;
    stage_step(G, Mo, dump, act, rec, ns - q, traj, stage_rows(sg_lds), S);
    if (!act) return;
    prep_step<F32, 1, StagedCells>(G, Mo, P, ns - q, rec, nullptr, dump, pkt + (size_t)g * PK, true, unit_scale(),
                                   nullptr, S);
  } else {
    if (!act) return;
    prep_step<F32>(G, Mo, P, ns - q, rec, nullptr, dump, pkt + (size_t)g * PK, true, unit_scale(), nullptr, NoVG(),
                   d32);
  }
}

/* Step packets of the first steps of the rays still being traced: n rays
 * (plist), np1 packets each, packet q of ray j (step np1 - q) at
 * pkt + (j np1 + q) PK.  Formed from the records the bounded phases stored,
 * while the unbounded phase extends the same trajectories past them. */
template <bool F32, bool STAGE>
__global__ __launch_bounds__(PREP_BLOCK) void k_prep_prefix(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P,
                                                            const int *__restrict__ plist, int n, int np1, int cap,
                                                            const double *__restrict__ traj,
                                                            const double *__restrict__ dump, double *__restrict__ pkt,
                                                            const float *__restrict__ d32)
{
  __shared__ double sg_lds[STAGE ? TPB * SG_WAVE_DBL : 1];
  const long long g = (long long)blockIdx.x * PREP_BLOCK + threadIdx.x;
  const bool act = g < (long long)n * np1;
  const int j = act ? (int)(g / np1) : 0, q = act ? (int)(g % np1) : 0;
  const double *rec = act ? traj + (size_t)plist[j] * cap * REC : nullptr;
  if constexpr (STAGE) {
    StagedCells S;
    stage_step(G, Mo, dump, act, rec, np1 - q, traj, stage_rows(sg_lds), S);
    if (!act) return;
    prep_step<F32, 1, StagedCells>(G, Mo, P, np1 - q, rec, nullptr, dump, pkt + (size_t)g * PK, true, unit_scale(),
                                   nullptr, S);
  } else {
    if (!act) return;
    prep_step<F32>(G, Mo, P, np1 - q, rec, nullptr, dump, pkt + (size_t)g * PK, true, unit_scale(), nullptr, NoVG(),
                   d32);
  }
}

/* The chains of the rays of a list, one lane per ray: SEG = false the rays
 * with seg_base[k] < 0 (recurrence over the packets), SEG = true those with
 * seg_base[k] >= 0 (composition of the segment maps). */
template <bool SEG>
__global__ __launch_bounds__(WAVE) void k_chain_list(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P,
                                                     const int *__restrict__ pix, const int *__restrict__ list, int n,
                                                     const long long *__restrict__ W, int cap, int stride,
                                                     const int *__restrict__ nstep_in,
                                                     const double *__restrict__ traj,
                                                     const long long *__restrict__ rbase,
                                                     const double *__restrict__ trajO,
                                                     const double *__restrict__ dump, const double *__restrict__ pkt,
                                                     const int *__restrict__ seg_base, int seg_len,
                                                     const double *__restrict__ segmap,
                                                     const int *__restrict__ segflag, double *__restrict__ out7,
                                                     int *__restrict__ flags, int *__restrict__ status,
                                                     unsigned long long *__restrict__ counts,
                                                     const int *__restrict__ pre_idx,
                                                     const double *__restrict__ pre_map,
                                                     const int *__restrict__ pre_flag, int npre)
{
  IPG_LATENCY_CRITICAL();
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int ray = list[k];
  int pi, pj;
  if (!ray_pixel(pix, ray, P.nx, P.ny, pi, pj)) return;
  const int ns = nstep_in[ray];
  double *o = out7 + (size_t)ray * 7;
  /* a ray past maxnstep is reported by each instance that sees it (a list
   * whose chains are all composed runs the SEG instance alone) */
  if (ns >= P.maxnstep - 1) {
    atomicMin(&status[1], pixel_code(pi, pj));
    for (int q = 0; q < 7; q++) o[q] = 0.;
    flags[ray] = 0;
    return;
  }
  const int sb = seg_base ? seg_base[k] : -1;
  if ((sb >= 0) != SEG) return;
  const double *rec = ray_rec(traj, cap, rbase, trajO, ray);
  double js, ks;
  if constexpr (SEG) {
    if (P.transfer_fp32) far_end_coefficients<true>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
    else far_end_coefficients<false>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
    IPG_SCHED_BARRIER();
  } else {
    if (P.transfer_fp32) far_end_coefficients<true>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
    else far_end_coefficients<false>(G, Mo, P, ns, rec, dump, js, ks, unit_scale());
  }
  double res[7];
  int nrad;
  if constexpr (SEG)
    flags[ray] = chain_ray_segmented(G, P, ns, seg_len, segmap + (size_t)sb * SEGCOL * NPAR, segflag + sb, js, ks,
                                     res, nrad, pre_idx ? pre_map + (size_t)pre_idx[k] * npre * SEGCOL * NPAR : nullptr,
                                     pre_idx ? pre_flag + (size_t)pre_idx[k] * npre : nullptr, pre_idx ? npre : 0);
  else
    flags[ray] = chain_ray<ILV>(G, P, ns, pkt + (size_t)W[k / ILV] * stride * ILV + pk_lane<ILV>(k % ILV), stride, js, ks,
                                res, nrad);
  for (int q = 0; q < 7; q++) o[q] = res[q];
  atomicAdd(&counts[1], (unsigned long long)nrad);
}

/* One lane per (segment, column): the composed maps of the long rays'
 * segments (columns 0..15 the N map of the basis states, 16 its source
 * term, 17 the unpolarized map).  Segments are given by the position of
 * their ray's first packet (in doubles) and their packets [q0, q1); PS is
 * the packets' field stride. */
template <int PS>
__global__ __launch_bounds__(WAVE) void k_seg(int nseg, const long long *__restrict__ seg_off,
                                              const int *__restrict__ seg_q0, const int *__restrict__ seg_q1,
                                              const double *__restrict__ pkt, double *__restrict__ segmap,
                                              int *__restrict__ segflag, int urgent)
{
  if (urgent) IPG_LATENCY_CRITICAL();
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nseg * SEGCOL) return;
  const int sgi = (int)(t / SEGCOL), b = (int)(t % SEGCOL);
  const int q0 = seg_q0[sgi], q1 = seg_q1[sgi];
  const double *pr = pkt + seg_off[sgi];
  double *o = segmap + ((size_t)sgi * SEGCOL + b) * NPAR;
  if (b == SEG_I) {
    double u[NPAR];
    segment_intensity<PS>(pr, q0, q1, u);
#pragma unroll
    for (int p = 0; p < NPAR; p++) o[p] = u[p];
    return;
  }
  Herm N;
  int fl = segment_column<PS>(pr, q0, q1, b, N);
#pragma unroll
  for (int p = 0; p < NPAR; p++) o[p] = herm_par(N, p);
  if (b == NPAR) segflag[sgi] = fl;
}

/* ------------------------------------------------------ adaptive refinement */

struct RefineK {
  int nxi, nyi, sp, nmin_y;       /* internal grid, base spacing, base columns */
  double refine_abs, refine_rel, refine_cut;
  int nearest;
};

/* Coarse-pass intensity estimate of src/main.c:579-632: the base rays'
 * Intensity (before the nu^3 scaling) weighted by the pixels each stands
 * for, summed in the coarse grid's order, times freqcgs^3 / (nx ny).  One
 * lane: the sum's order is the reference's. */
__global__ void k_refine_iavg(RefineK R, int nbase, int full, double f3, const double *__restrict__ out7,
                              double *__restrict__ iavg)
{
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const size_t sp = (size_t)R.sp, nx = (size_t)R.nxi, ny = (size_t)R.nyi;
  double tot = 0.;
  for (int r = 0; r < nbase; r++) {
    const size_t i = (size_t)(r / R.nmin_y) * sp, j = (size_t)(r % R.nmin_y) * sp;
    /* the ray of base pixel r: r itself in the coarse pass's list, or the
     * pixel's place in a render of the whole grid (full) */
    const size_t ray = full ? i * ny + j : (size_t)r;
    const double I = out7[ray * 7];
    double w;
    if (i % (nx - 1) != 0 && j % (ny - 1) != 0) w = I * sp * sp;
    else if ((i == 0 && j % (ny - 1) != 0) || (i % (nx - 1) != 0 && j == 0)) w = I * (sp / 2 + 1) * sp;
    else if (i % (nx - 1) != 0 || j % (ny - 1) != 0) w = I * (sp / 2) * sp;
    else if (i == 0 && j == 0) w = I * (sp / 2 + 1) * (sp / 2 + 1);
    else if (i == 0 || j == 0) w = I * (sp / 2 + 1) * sp / 2;
    else w = I * (sp * sp) / 4;
    tot += w;
  }
  iavg[0] = tot * f3 / (double)(nx * ny);
}

/* One refinement level (src/main.c:635-776) at spacing ns: each pixel of the
 * level not on the coarser grid either joins the list of rays to trace
 * (both error estimates above their tolerances and the lower neighbour
 * above refine_cut) or is interpolated from the coarser grid at once
 * (save_pixelTransfer, lininterp2, lininterp4: src/main.c:912-1000).  The
 * pixels a level reads all lie on the coarser grid, so its pixels are
 * independent.  count[0]: traced, count[1]: interpolated.  With list null
 * every pixel has been traced beforehand, and those refined keep their
 * values. */
__global__ void k_refine_level(RefineK R, int ns, int unpol, const double *__restrict__ iavg,
                               double *__restrict__ image, double *__restrict__ imageS,
                               double *__restrict__ taus, int *__restrict__ nstep_pix,
                               int *__restrict__ flag_pix, int *__restrict__ interp, int *__restrict__ list,
                               unsigned int *__restrict__ count)
{
  const int na = (R.nxi - 1) / ns + 1, nb = (R.nyi - 1) / ns + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)na * nb) return;
  const int a = (int)(t / nb), b = (int)(t % nb);
  const bool ca = (a & 1) == 0, cb = (b & 1) == 0;
  if (ca && cb) return;                                  /* traced before */
  const int i = a * ns, j = b * ns;
  const size_t ny = (size_t)R.nyi;
  auto at = [&](int ii, int jj) { return (size_t)ii * ny + jj; };
  const double Iavg = iavg[0];
  size_t src[4];
  int nsrc;
  double I1, err_abs, err_rel;
  if (ca) {                                              /* on a coarser column */
    src[0] = at(i, j - ns);
    src[1] = at(i, j + ns);
    nsrc = 2;
    I1 = image[src[0]];
    const double I2 = image[src[1]];
    err_abs = (I2 - I1) / 2 / Iavg;
    err_rel = (I2 - I1) / 2 / I1;
  } else if (cb) {                                       /* on a coarser row */
    src[0] = at(i - ns, j);
    src[1] = at(i + ns, j);
    nsrc = 2;
    I1 = image[src[0]];
    const double I2 = image[src[1]];
    err_abs = (I2 - I1) / 2 / Iavg;
    err_rel = (I2 - I1) / 2 / I1;
  } else {                                               /* between four corners */
    src[0] = at(i - ns, j - ns);
    src[1] = at(i + ns, j - ns);
    src[2] = at(i - ns, j + ns);
    src[3] = at(i + ns, j + ns);
    nsrc = 4;
    I1 = image[src[0]];
    const double I2 = image[src[1]], I3 = image[src[2]];
    err_abs = ((I2 + I3) / 2 - I1) / Iavg;
    err_rel = (I2 + I3) / (2 * I1) - 1.;
  }
  const size_t p = at(i, j);
  if ((fabs(err_abs) > R.refine_abs && fabs(err_rel) > R.refine_rel) && fabs(I1) > R.refine_cut) {
    const unsigned slot = atomicAdd(&count[0], 1u);
    if (list) list[slot] = (i << 16) | j;   /* no list: the pixel was traced already */
    interp[p] = 0;
    return;
  }
  atomicAdd(&count[1], 1u);
  interp[p] = 1;
  nstep_pix[p] = 0;
  flag_pix[p] = 0;
  if (R.nearest) {
    const size_t q = src[0];
    image[p] = image[q];
    if (!unpol) {
      taus[p] = taus[q];
      for (int k = 0; k < 5; k++) imageS[p * 5 + k] = imageS[q * 5 + k];
    } else {
      taus[p] = 0.;   /* save_pixelTransfer leaves tau as allocated (0) with only_intensity */
    }
  } else if (nsrc == 2) {
    image[p] = .5 * (image[src[0]] + image[src[1]]);
    taus[p] = .5 * (taus[src[0]] + taus[src[1]]);
    if (!unpol)
      for (int k = 0; k < 5; k++) imageS[p * 5 + k] = .5 * (imageS[src[0] * 5 + k] + imageS[src[1] * 5 + k]);
  } else {
    image[p] = .25 * (image[src[0]] + image[src[1]] + image[src[2]] + image[src[3]]);
    taus[p] = .25 * (taus[src[0]] + taus[src[1]] + taus[src[2]] + taus[src[3]]);
    if (!unpol)
      for (int k = 0; k < 5; k++)
        imageS[p * 5 + k] = .25 * (imageS[src[0] * 5 + k] + imageS[src[1] * 5 + k] + imageS[src[2] * 5 + k]
                                   + imageS[src[3] * 5 + k]);
  }
}

/* ------------------------------------------------------- pixel trace output */

/* One lane works (trace_pixel_records): a debugging and parity tool. */
__global__ void k_trace_pixel(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P, int pi, int pj, int cap, int prims_ok,
                              const double *__restrict__ dump, double *__restrict__ rec, int *__restrict__ nt,
                              double *__restrict__ out, int *__restrict__ nstep_out)
{
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  trace_pixel_records(G, Mo, P, pi, pj, cap, prims_ok, dump, rec, nt, out, nstep_out);
}

/* ------------------------------------ single-image render without allocation
 *
 * ipgpu_render_noalloc works in one fixed scratch array per GPU: status
 * words at byte 0, counters at byte 64, per-ray step counts and record
 * offsets, then an arena of ray records.  A count pass gives every ray's
 * step count; the prefix sum S_r = off[r] + r places ray r's nstep + 1
 * records in a virtual contiguous layout, and chunk c holds the rays with
 * S_r in [c A', (c+1) A'), A' = arena records - (maxnstep + 1), so that every
 * ray of a chunk fits the arena at S_r - c A'.  Per chunk the rays are
 * traced again into the arena and transferred (transfer_fused) straight into
 * the caller's output arrays. */

/* status[1], status[2] (first maxnstep and NaN pixel codes) to "none",
 * status[0] (store passes whose step count differs) and the counters to 0 */
__global__ void k_na_init(int *__restrict__ status, unsigned long long *__restrict__ counts)
{
  if (threadIdx.x == 0) {
    status[0] = 0;
    status[1] = 0x7fffffff;
    status[2] = 0x7fffffff;
    status[3] = 0;
    counts[0] = 0;
    counts[1] = 0;
  }
}

/* Count pass on lane pairs: every ray's step count, nothing stored. */
__global__ __launch_bounds__(WAVE) void k_na_count(ipgpu_geom G, ipgpu_image P, int nrays, int *__restrict__ nstep_out,
                                                   unsigned long long *__restrict__ counts)
{
  IPG_LATENCY_CRITICAL();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ray = t >> 1;
  const bool odd = t & 1;
  if (ray >= nrays) return;
  int pi, pj, ns = 0;
  if (ray_pixel(nullptr, ray, P.nx, P.ny, pi, pj)) ns = trace_pair(G, P, pi, pj, 0, nullptr, 0, 0x7fffffff, nullptr, odd);
  if (odd) return;
  nstep_out[ray] = ns;
  atomicAdd(&counts[0], (unsigned long long)ns);
}

/* Store pass of chunk c on lane pairs: the records of its rays at
 * arena + (S_r - c A') REC; a step count other than the count pass's is
 * reported in status[0]. */
__global__ __launch_bounds__(WAVE) void k_na_store(ipgpu_geom G, ipgpu_image P, int nrays, long long c, long long Ap,
                                                   const int *__restrict__ nstep_in, const long long *__restrict__ off,
                                                   double *__restrict__ arena, int *__restrict__ status)
{
  IPG_LATENCY_CRITICAL();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ray = t >> 1;
  const bool odd = t & 1;
  if (ray >= nrays) return;
  const long long S = off[ray] + ray;
  if (S / Ap != c) return;
  int pi, pj;
  if (!ray_pixel(nullptr, ray, P.nx, P.ny, pi, pj)) return;
  const int ns = nstep_in[ray];
  if (ns >= P.maxnstep - 1) return;
  const int n = trace_pair(G, P, pi, pj, ns + 1, arena + (size_t)(S - c * Ap) * REC, 0, 0x7fffffff, nullptr, odd);
  if (!odd && n != ns) atomicAdd(&status[0], 1);
}

/* Transfer of chunk c, one lane per ray, into the caller's arrays in the
 * layout of save_pixel (src/main.c:885-910); any output may be null. */
template <bool F32>
__global__ __launch_bounds__(WAVE) void k_na_xfer(ipgpu_geom G, ipgpu_model Mo, ipgpu_image P, int nrays, long long c,
                                                  long long Ap, const int *__restrict__ nstep_in,
                                                  const long long *__restrict__ off, const double *__restrict__ arena,
                                                  const double *__restrict__ dump, double *__restrict__ image,
                                                  double *__restrict__ imageS, double *__restrict__ taus,
                                                  int *__restrict__ nstep_pix, int *__restrict__ flag_pix,
                                                  int *__restrict__ status, unsigned long long *__restrict__ counts)
{
  const int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= nrays) return;
  const long long S = off[ray] + ray;
  if (S / Ap != c) return;
  int pi, pj;
  if (!ray_pixel(nullptr, ray, P.nx, P.ny, pi, pj)) return;
  const size_t p = (size_t)pi * P.ny + pj;
  const int ns = nstep_in[ray];
  if (nstep_pix) nstep_pix[p] = ns;
  double o[7] = {0., 0., 0., 0., 0., 0., 0.};
  int flag = 0, nrad = 0;
  __shared__ double mi_lds[F32 ? 1 : WAVE * 17];
  if (ns >= P.maxnstep - 1) atomicMin(&status[1], pixel_code(pi, pj));
  else
    flag = transfer_fused<F32>(G, Mo, P, ns, arena + (size_t)(S - c * Ap) * REC, nullptr, dump, unit_scale(), o, nrad,
                               F32 ? nullptr : mi_lds + 17 * threadIdx.x);
  const double f3 = P.freqcgs * P.freqcgs * P.freqcgs;
  if (image) image[p] = o[0] * f3;
  if (taus) taus[p] = o[5];
  if (flag_pix) flag_pix[p] = flag;
  if (!P.only_unpolarized) {
    const double qs = (P.qu_conv == 0) ? -1. : 1.;
    if (imageS) {
      imageS[p * 5 + 0] = o[1] * f3;
      imageS[p * 5 + 1] = qs * o[2] * f3;
      imageS[p * 5 + 2] = qs * o[3] * f3;
      imageS[p * 5 + 3] = o[4] * f3;
      imageS[p * 5 + 4] = o[6];
    }
    if (isnan(o[1])) atomicMin(&status[2], pixel_code(pi, pj));
  }
  atomicAdd(&counts[1], (unsigned long long)nrad);
}

}  // namespace

/* The working memory of ipgpu_render_noalloc (see above), one copy in each
 * GPU's memory for as long as the process runs.  Host code never uses this
 * name as an address: the device address comes from hipGetSymbolAddress,
 * once per device (noalloc_setup), and calls on one GPU share the array. */
static __device__ __align__(256) unsigned char scratch[IPGPU_SCRATCH_BYTES];

/* ------------------------------------------------------------------ C-ABI */

constexpr int NEV = 24;
constexpr int MAXPH = 5;   /* trace phases of the streamed render (the last unbounded) */

struct ipgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t streamF[MAXPH - 1] = {}; /* transfer of the rays finished in each bounded phase */
  hipEvent_t ev[NEV] = {};
  char err[512] = {0};
  ipgpu_geom geom{};
  bool has_geom = false;
  ipgpu_model model{};
  bool has_model = false;
  double *d_dump = nullptr;
  float *d_dump32 = nullptr;        /* FP32 copy of the records (mixed-precision transfers) */
  /* the batch's first-pass overflow: arena chunks for the records past the
   * slot, their table, a per-ray flag and the chunk counter */
  double *d_ovf_arena = nullptr;
  size_t ovf_arena_doubles = 0;
  int *d_ovf_table = nullptr, *d_ovf_ok = nullptr, *d_glist = nullptr;
  size_t cap_ovf_table = 0, cap_ovf_ok = 0, cap_glist = 0;
  long long *d_goff2 = nullptr;
  size_t cap_goff2 = 0;
  unsigned long long *d_ovf_cnt = nullptr;
  size_t cap_ovf_cnt = 0;
  size_t ovf_arena_bytes = (size_t)8 << 30;
  const float *r_d32 = nullptr;     /* the copy the single-image render in progress reads, or null */
  size_t dump32_cells = 0;
  size_t dump_cells = 0;
  double *d_traj = nullptr;
  size_t traj_doubles = 0;
  int *d_nturns = nullptr;
  /* batches: records per ray of the first trace pass; the full-length
   * records of the rays longer than that */
  int batch_slot = 2048;
  int2 *h_items = nullptr;     /* pinned (job, ray) items of the in-place path */
  size_t h_items_cap = 0;
  std::vector<int> h_gid;
  double *d_traj2 = nullptr;
  size_t traj2_doubles = 0;
  long long *d_off2 = nullptr;
  size_t cap_off2 = 0;
  size_t nturns_ints = 0;
  double *d_pkt = nullptr;
  size_t pkt_doubles = 0;
  double *d_aux = nullptr;
  size_t aux_doubles = 0;
  long long *d_off = nullptr;
  size_t cap_off = 0;
  int *d_blk = nullptr;
  size_t cap_blk = 0;
  int *d_segbase = nullptr;
  size_t cap_segbase = 0;
  int *d_nstep_ray = nullptr, *d_flag_ray = nullptr;
  double *d_out7 = nullptr;
  double *d_image = nullptr, *d_imageS = nullptr, *d_taus = nullptr;
  int *d_nstep_pix = nullptr, *d_flag_pix = nullptr;
  size_t cap_nstep_ray = 0, cap_flag_ray = 0, cap_out7 = 0, cap_image = 0, cap_taus = 0;
  size_t cap_nstep_pix = 0, cap_flag_pix = 0, cap_imageS = 0;
  int *d_status = nullptr;
  unsigned long long *d_counts = nullptr;
  unsigned long long counts[2] = {0, 0};
  int last_npix = 0, last_unpol = 0, have_result = 0;
  float ms_geo = 0, ms_xfer = 0, ms_total = 0, ms_stage[4] = {0, 0, 0, 0};
  int bad_i = -1, bad_j = -1, bad_job = -1;
  /* records per ray of a single-image render: a hint (the longest ray of
   * the last render, at least 8192, within the render memory); rays past it
   * continue in full-length regions (d_trajO, offsets per ray d_rbase) */
  int traj_cap_steps = 8192;
  double *d_trajO = nullptr;
  size_t trajO_doubles = 0;
  int *d_ovl = nullptr;
  long long *d_ovoff = nullptr, *d_rbase = nullptr;
  size_t cap_ovl = 0, cap_ovoff = 0, cap_rbase = 0;
  std::vector<int> h_ovl;
  std::vector<long long> h_ovoff;
  int last_max_steps = 0, last_overflow = 0;
  /* the streamed render's record pool: each ray still live when the last
   * bounded phase ends gets a full-length region of d_trajO, assigned on
   * the device (k_pool_assign), so the per-ray slots need hold only the
   * bounded phases' records */
  int pool_on = 1;
  int pool_max = -1;                /* at most this many regions (-1: no limit; 0: no pool) */
  double live_frac = 1.0 / 16;      /* rays live after the last bounded phase, of the last render */
  unsigned int *d_pool_ctr = nullptr;
  int *d_pool_list = nullptr;
  size_t cap_pool_list = 0, cap_pool_ctr = 0;
  int last_pool_regions = 0, last_pool_used = 0;
  /* device memory a single-image render may use for its ray records and
   * step packets; larger images are rendered in chunks of rays */
  size_t render_bytes = (size_t)64 << 30;
  double avg_steps_hint = 2048.;    /* mean step count per ray, from the last render */
  double pkt_bytes_per_ray = 0.;    /* step packets per ray of the last streamed render */
  int last_chunks = 1;
  std::vector<int> h_pixlist;
  int *d_pixlist = nullptr;
  size_t cap_pixlist = 0;
  size_t dev_bytes = 0;             /* device memory held by the context */
  void *d_scratch = nullptr;        /* this GPU's scratch (ipgpu_render_noalloc) */
  /* streamed render: step budget of the last bounded trace phase (0: one
   * pass), the ray length above which a chain is composed from segments,
   * the segment length */
  int trace_budget = 3584;
  int trace_phases = 2;      /* bounded phases, budgets trace_budget j / trace_phases */
  int seg_min_steps = SEG_MIN_STEPS;
  int seg_len = SEG_LEN;
  int *d_live = nullptr, *d_list = nullptr, *d_blk2 = nullptr;
  long long *d_offl = nullptr;
  size_t cap_live = 0, cap_list = 0, cap_blk2 = 0, cap_offl = 0;
  double *d_pkt2 = nullptr;
  size_t pkt2_doubles = 0;
  long long last_total = 0;  /* packets of the previous render (arena size hint) */
  /* streamed render: the first steps of the rays still traced after the
   * last bounded phase are prepared and composed while the unbounded phase
   * runs (packets d_ppkt, list d_plist, list positions d_pidx) */
  int tail_prefix = 1;
  int gather_stage = 0;   /* packet kernels stage their dump gathers in LDS */
  int *d_wtile = nullptr;  /* a batch's first-pass tiles in dispatch order */
  size_t cap_wtile = 0;
  double *d_ppkt = nullptr;
  size_t ppkt_doubles = 0;
  int *d_plist = nullptr, *d_pidx = nullptr;
  size_t cap_plist = 0, cap_pidx = 0;
  std::vector<int> h_ppos;
  struct SegSet {
    long long *d_off = nullptr;
    int *d_q0 = nullptr, *d_q1 = nullptr, *d_flag = nullptr;
    double *d_map = nullptr;
    size_t cap_off = 0, cap_q0 = 0, cap_q1 = 0, cap_flag = 0, cap_map = 0;
    std::vector<long long> off;
    std::vector<int> q0, q1;
  } segset[MAXPH], segpre;
  /* host copies read by pending asynchronous copies: they live until the
   * render that filled them has synchronised */
  std::vector<int> h_ns, h_list, h_base, h_cnt, h_tmp;
  std::vector<long long> h_off, h_eoff;
  /* pinned host memory of the streamed render: per phase the step counts
   * and live flags, then lists, packet offsets and segment bases */
  void *h_pin = nullptr;
  size_t h_pin_bytes = 0;
  int streamed = 0;
  int nphase = 0;
  /* batches */
  BGroup *d_groups = nullptr;
  size_t cap_groups = 0;
  BJob *d_jobs = nullptr;
  size_t cap_jobs = 0;
  int2 *d_items = nullptr;
  size_t cap_items = 0;
  double *d_bimage = nullptr, *d_bimageS = nullptr, *d_btaus = nullptr;
  int *d_bnstep = nullptr, *d_bflag = nullptr;
  size_t cap_bimage = 0, cap_bimageS = 0, cap_btaus = 0, cap_bnstep = 0, cap_bflag = 0;
  unsigned long long *d_bstatus = nullptr;
  size_t batch_traj_bytes = (size_t)160 << 30; /* geodesic records per pass of a batch */
  /* the batch transfer: persistent work-queue launch (1) or one wave per 64
   * items (0); persistent waves (0: 4 per compute unit) */
  int batch_queue = 1;
  int batch_queue_waves = 0;
  /* the transfer of up to MJ jobs of a trace group per lane (shared
   * geometry), or one job per lane */
  int batch_multi = 1;
  int4 *d_subsets = nullptr;
  size_t cap_subsets = 0;
  /* the last batch's work: items, their transfer steps and radiating steps
   * (a multi-job item counts once), (job, ray) pairs and their steps */
  unsigned long long *d_bcounts = nullptr;
  long long bstat[5] = {0, 0, 0, 0, 0};
  int ncu = 256;
  int xq_occ[2] = {1, 1};   /* resident waves per compute unit of k_xfer_queue<FP64>, <F32> */
  int xm_occ[2] = {1, 1};   /* the same of k_xfer_multi */
  unsigned int *d_queue = nullptr;
  /* the last in-place batch, for ipgpu_rerender_batch: its images, trace
   * scales, output offsets, items and the graph of its transfer (job table
   * upload from pinned memory, queue reset, transfer) */
  struct LastBatch {
    bool valid = false;
    int njobs = 0, nitems = 0, f32 = 0, multi = 0;
    std::vector<ipgpu_image> img;
    std::vector<BJob> bj;
    long long pix_total = 0;
    const double *traj2 = nullptr;
  } blast;
  /* what the captured transfer graph baked in: kernel arguments by value
   * and the dump it reads */
  struct GraphKey {
    ipgpu_geom geom;
    ipgpu_model model;
    const double *dump;
    const float *dump32;
    int queue, waves;
  } bkey;
  BJob *h_bjobs = nullptr;          /* pinned job table the graph uploads */
  size_t h_bjobs_cap = 0;
  hipGraph_t bgraph = nullptr;
  hipGraphExec_t bgraph_exec = nullptr;
  int rerender_graph = 1;           /* replay the transfer from a captured graph */
  /* adjoint */
  ipgpu_units units{};
  bool has_units = false;
  Herm *d_ckpt = nullptr, *d_seg = nullptr;
  size_t cap_ckpt = 0, cap_seg = 0;
  /* adjoint in segments: steps per segment, segment and lane tables, maps,
   * states before and adjoint states after each segment, per-segment sums */
  int adj_seg_len = 512;
  std::vector<int> h_aseg, h_alane, h_aray;
  int *d_aint = nullptr;
  double *d_amap = nullptr, *d_ag3 = nullptr;
  Herm *d_aN = nullptr, *d_aL = nullptr;
  size_t cap_aint = 0, cap_amap = 0, cap_ag3 = 0, cap_aN = 0, cap_aL = 0;
  double *d_gcell = nullptr, *d_out4 = nullptr, *d_target = nullptr, *d_scale = nullptr;
  /* the pixel list of the rays being rendered (nullptr: the image's tiles)
   * and the row length of the output image they are written to */
  const int *pix = nullptr;
  int out_ny = 0;
  /* adaptive refinement: pixel list, decision counters, per-level counts */
  int *d_rlist = nullptr;
  size_t cap_rlist = 0;
  unsigned long long *d_rcount = nullptr;
  double *d_iavg = nullptr;
  int *d_interp = nullptr;
  size_t cap_interp = 0;
  int ref_levels = 0;
  long long refine_spec_max = 1ll << 20;   /* speculative refinement up to this many internal pixels */
  double ref_iavg = 0.;
  long long ref_traced[IPGPU_MAX_LEVELS] = {0}, ref_interp[IPGPU_MAX_LEVELS] = {0};
  /* pixel trace output */
  double *d_trrec = nullptr, *d_trout = nullptr;
  int *d_trnt = nullptr;
  size_t cap_trrec = 0, cap_trout = 0, cap_trnt = 0;
  size_t cap_gcell = 0, cap_out4 = 0, cap_target = 0, cap_scale = 0;
};

#define HIPCHK(ctx, call)                                                                    \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      snprintf((ctx)->err, sizeof((ctx)->err), "%s failed: %s", #call, hipGetErrorString(e_)); \
      return IPGPU_ERR_HIP;                                                                  \
    }                                                                                        \
  } while (0)

static int set_err(ipgpu_ctx *ctx, int code, const char *msg)
{
  if (ctx) snprintf(ctx->err, sizeof(ctx->err), "%s", msg);
  return code;
}

template <typename T>
static int grow(ipgpu_ctx *ctx, T *&ptr, size_t &cap, size_t need)
{
  if (need <= cap && ptr) return 0;
  if (ptr) {
    HIPCHK(ctx, hipFree(ptr));
    ctx->dev_bytes -= cap * sizeof(T);
  }
  ptr = nullptr;
  cap = 0;
  hipError_t e = hipMalloc((void **)&ptr, need * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    snprintf(ctx->err, sizeof(ctx->err), "device allocation of %zu bytes failed (%s); the context holds %zu bytes",
             need * sizeof(T), hipGetErrorString(e), ctx->dev_bytes);
    return e == hipErrorOutOfMemory ? IPGPU_ERR_NOMEM : IPGPU_ERR_HIP;
  }
  cap = need;
  ctx->dev_bytes += need * sizeof(T);
  return 0;
}

/* grow that keeps the first `keep` elements (copied on the stream, which
 * must hold no other work on the buffer in flight). */
template <typename T>
static int grow_keep(ipgpu_ctx *ctx, T *&ptr, size_t &cap, size_t need, size_t keep, hipStream_t st)
{
  if (need <= cap && ptr) return 0;
  if (!ptr || keep == 0) return grow(ctx, ptr, cap, need);
  T *np = nullptr;
  hipError_t e = hipMalloc((void **)&np, need * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    snprintf(ctx->err, sizeof(ctx->err), "device allocation of %zu bytes failed (%s); the context holds %zu bytes",
             need * sizeof(T), hipGetErrorString(e), ctx->dev_bytes);
    return e == hipErrorOutOfMemory ? IPGPU_ERR_NOMEM : IPGPU_ERR_HIP;
  }
  HIPCHK(ctx, hipMemcpyAsync(np, ptr, std::min(keep, cap) * sizeof(T), hipMemcpyDeviceToDevice, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  HIPCHK(ctx, hipFree(ptr));
  ctx->dev_bytes += (need - cap) * sizeof(T);
  ptr = np;
  cap = need;
  return 0;
}

/* grow for the streamed render's arenas: when it must reallocate, a quarter
 * more than asked (as far as 95% of the render memory leaves room), so that the
 * next renders, whose sizes follow this one's, do not reallocate again (a
 * reallocation of tens of GB stalls the render for hundreds of ms). */
template <typename T>
static int grow_hr(ipgpu_ctx *ctx, T *&ptr, size_t &cap, size_t need)
{
  if (need <= cap && ptr) return 0;
  const size_t held = (ctx->traj_doubles + ctx->pkt_doubles + ctx->pkt2_doubles + ctx->ppkt_doubles
                       + ctx->trajO_doubles) * sizeof(double) - cap * sizeof(T) + need * sizeof(T);
  size_t extra = need / 4;
  const size_t lim = ctx->render_bytes / 20 * 19;   /* 95%: the rest for per-ray tables and maps */
  const size_t room = lim > held ? (lim - held) / sizeof(T) : 0;
  if (extra > room) extra = room;
  return grow(ctx, ptr, cap, need + extra);
}

/* Device address of scratch on each GPU, looked up once per device with
 * that device current. */
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch[MAX_DEVICES] = {};

static int noalloc_setup(int device, void **ptr)
{
  if (device < 0 || device >= MAX_DEVICES) return IPGPU_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return IPGPU_ERR_HIP;
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  return IPGPU_OK;
}

extern "C" {

int ipgpu_create(ipgpu_ctx **out, int device)
{
  if (!out) return IPGPU_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return IPGPU_ERR_HIP;
  if (device < 0 || device >= ndev) return IPGPU_ERR_ARG;
  ipgpu_ctx *ctx = new ipgpu_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return IPGPU_ERR_HIP;
  }
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    ctx->ncu = ncu;
  {
    /* single-image render memory: 60% of the device's (173 GB of an
     * MI355X's 288), so that the benchmarked images run whole */
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && totalb > ((size_t)4 << 30))
      ctx->render_bytes = (size_t)(0.6 * (double)totalb);
  }
  for (int f = 0; f < 2; f++) {
    int nb = 0;
    hipError_t e = f ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_xfer_queue<true>, WAVE, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_xfer_queue<false>, WAVE, 0);
    if (e == hipSuccess && nb > 0) ctx->xq_occ[f] = nb;
    nb = 0;
    e = f ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_xfer_multi<true>, WAVE, 0)
          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_xfer_multi<false>, WAVE, 0);
    if (e == hipSuccess && nb > 0) ctx->xm_occ[f] = nb;
  }
  if (const char *e = getenv("IPGPU_BATCH_QUEUE")) ctx->batch_queue = atoi(e);
  if (const char *e = getenv("IPGPU_RECORD_POOL_MAX")) ctx->pool_max = atoi(e);
  if (const char *e = getenv("IPGPU_BATCH_MULTI")) ctx->batch_multi = atoi(e);
  if (const char *e = getenv("IPGPU_BATCH_QUEUE_WAVES")) ctx->batch_queue_waves = atoi(e);
  if (noalloc_setup(device, &ctx->d_scratch) != IPGPU_OK) {
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return IPGPU_ERR_HIP;
  }
  for (auto &st : ctx->streamF)
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      return IPGPU_ERR_HIP;
    }
  for (int q = 0; q < NEV; q++)
    if (hipEventCreate(&ctx->ev[q]) != hipSuccess) {
      delete ctx;
      return IPGPU_ERR_HIP;
    }
  if (const char *e = getenv("IPGPU_TRACE_BUDGET")) ctx->trace_budget = atoi(e);
  if (const char *e = getenv("IPGPU_SEG_MIN_STEPS")) ctx->seg_min_steps = atoi(e);
  if (const char *e = getenv("IPGPU_SEG_LEN")) ctx->seg_len = atoi(e) > 1 ? atoi(e) : 1;
  if (const char *e = getenv("IPGPU_TRACE_PHASES")) ctx->trace_phases = atoi(e);
  if (const char *e = getenv("IPGPU_REFINE_SPECULATE_MAX")) ctx->refine_spec_max = atoll(e);
  if (const char *e = getenv("IPGPU_TAIL_PREFIX")) ctx->tail_prefix = atoi(e);
  if (const char *e = getenv("IPGPU_BATCH_SLOT")) ctx->batch_slot = atoi(e);
  if (const char *e = getenv("IPGPU_BATCH_ARENA_MIB")) ctx->ovf_arena_bytes = (size_t)std::max(0, atoi(e)) << 20;
  if (const char *e = getenv("IPGPU_ADJ_SEG_LEN")) ctx->adj_seg_len = atoi(e) > 0 ? atoi(e) : 512;
  if (hipMalloc((void **)&ctx->d_status, NSTATUS * sizeof(int)) != hipSuccess
      || hipMalloc((void **)&ctx->d_counts, 2 * sizeof(unsigned long long)) != hipSuccess) {
    delete ctx;
    return IPGPU_ERR_HIP;
  }
  *out = ctx;
  return IPGPU_OK;
}

void ipgpu_destroy(ipgpu_ctx *ctx)
{
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (auto st : ctx->streamF)
    if (st) hipStreamSynchronize(st);
  if (ctx->h_pin) hipHostFree(ctx->h_pin);
  if (ctx->h_items) hipHostFree(ctx->h_items);
  if (ctx->h_bjobs) hipHostFree(ctx->h_bjobs);
  if (ctx->bgraph_exec) hipGraphExecDestroy(ctx->bgraph_exec);
  if (ctx->bgraph) hipGraphDestroy(ctx->bgraph);
  for (auto *S : {&ctx->segset[0], &ctx->segset[1], &ctx->segset[2], &ctx->segset[3], &ctx->segset[4],
                  &ctx->segpre}) {
    void *sp[] = {S->d_off, S->d_q0, S->d_q1, S->d_flag, S->d_map};
    for (void *p : sp)
      if (p) hipFree(p);
  }
  void *ptrs[] = {ctx->d_ovf_arena, ctx->d_ovf_table, ctx->d_ovf_ok, ctx->d_glist, ctx->d_goff2, ctx->d_ovf_cnt,
                  ctx->d_live, ctx->d_list, ctx->d_blk2, ctx->d_offl, ctx->d_pkt2, ctx->d_dump, ctx->d_dump32, ctx->d_traj, ctx->d_pkt, ctx->d_aux, ctx->d_off, ctx->d_blk, ctx->d_segbase, ctx->d_nturns, ctx->d_nstep_ray, ctx->d_flag_ray,
                  ctx->d_out7, ctx->d_image, ctx->d_imageS, ctx->d_taus, ctx->d_nstep_pix,
                  ctx->d_flag_pix, ctx->d_status, ctx->d_counts,
                  ctx->d_groups, ctx->d_jobs, ctx->d_items, ctx->d_bimage, ctx->d_bimageS, ctx->d_btaus,
                  ctx->d_bnstep, ctx->d_bflag, ctx->d_bstatus, ctx->d_ckpt, ctx->d_seg, ctx->d_gcell,
                  ctx->d_out4, ctx->d_target, ctx->d_scale, ctx->d_trrec, ctx->d_trout, ctx->d_trnt,
                  ctx->d_rlist, ctx->d_rcount, ctx->d_iavg, ctx->d_interp, ctx->d_ppkt, ctx->d_plist, ctx->d_traj2, ctx->d_off2,
                  ctx->d_pidx, ctx->d_trajO, ctx->d_ovl, ctx->d_ovoff, ctx->d_rbase, ctx->d_pixlist,
                  ctx->d_aint, ctx->d_amap, ctx->d_ag3, ctx->d_aN, ctx->d_aL, ctx->d_queue, ctx->d_pool_ctr,
                  ctx->d_pool_list, ctx->d_subsets, ctx->d_bcounts, ctx->d_wtile};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (int q = 0; q < NEV; q++)
    if (ctx->ev[q]) hipEventDestroy(ctx->ev[q]);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  for (auto st : ctx->streamF)
    if (st) hipStreamDestroy(st);
  delete ctx;
}

const char *ipgpu_last_error(const ipgpu_ctx *ctx) { return ctx ? ctx->err : "null context"; }

void *ipgpu_stream(ipgpu_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

static void batch_evict(ipgpu_ctx *ctx);
static void release_overflow(ipgpu_ctx *ctx);

int ipgpu_set_geometry(ipgpu_ctx *ctx, const ipgpu_geom *g)
{
  if (!ctx || !g) return IPGPU_ERR_ARG;
  if (g->metric != IPGPU_METRIC_MKS && g->metric != IPGPU_METRIC_FMKS
      && g->metric != IPGPU_METRIC_EMINKOWSKI)
    return set_err(ctx, IPGPU_ERR_ARG, "unsupported metric (MKS, FMKS, eMinkowski)");
  if (g->N1 <= 0 || g->N2 <= 0 || g->N3 <= 0) return set_err(ctx, IPGPU_ERR_ARG, "bad grid size");
  batch_evict(ctx);
  ctx->geom = *g;
  ctx->has_geom = true;
  return IPGPU_OK;
}

int ipgpu_set_model(ipgpu_ctx *ctx, const ipgpu_model *m)
{
  if (!ctx || !m) return IPGPU_ERR_ARG;
  if (m->kind != IPGPU_MODEL_IHARM && m->kind != IPGPU_MODEL_SPHERE)
    return set_err(ctx, IPGPU_ERR_ARG, "unsupported model kind");
  ctx->model = *m;
  ctx->has_model = true;
  return IPGPU_OK;
}

static size_t ghosted_cells(const ipgpu_geom &g)
{
  return (size_t)(g.N1 + 2) * (g.N2 + 2) * (g.N3 + 2);
}

static int ensure_dump(ipgpu_ctx *ctx)
{
  size_t nc = ghosted_cells(ctx->geom);
  if (ctx->d_dump && ctx->dump_cells == nc) return 0;
  if (ctx->d_dump) {
    HIPCHK(ctx, hipFree(ctx->d_dump));
    ctx->dev_bytes -= ctx->dump_cells * CELL_STRIDE * sizeof(double);
  }
  ctx->d_dump = nullptr;
  ctx->dump_cells = 0;
  HIPCHK(ctx, hipMalloc((void **)&ctx->d_dump, nc * CELL_STRIDE * sizeof(double)));
  ctx->dump_cells = nc;
  ctx->dev_bytes += nc * CELL_STRIDE * sizeof(double);
  return 0;
}

/* The FP32 copy of the cell records for a mixed-precision batch (its
 * contents are written before each transfer launch that reads it). */
static int ensure_dump32(ipgpu_ctx *ctx)
{
  if (!ctx->d_dump || ctx->model.kind != IPGPU_MODEL_IHARM) return 0;
  if (ctx->d_dump32 && ctx->dump32_cells == ctx->dump_cells) return 0;
  if (ctx->d_dump32) {
    HIPCHK(ctx, hipFree(ctx->d_dump32));
    ctx->dev_bytes -= ctx->dump32_cells * CELL32 * sizeof(float);
  }
  ctx->d_dump32 = nullptr;
  ctx->dump32_cells = 0;
  HIPCHK(ctx, hipMalloc((void **)&ctx->d_dump32, ctx->dump_cells * CELL32 * sizeof(float)));
  ctx->dump32_cells = ctx->dump_cells;
  ctx->dev_bytes += ctx->dump_cells * CELL32 * sizeof(float);
  return 0;
}

int ipgpu_upload_fields(ipgpu_ctx *ctx, const double *const fields[IPGPU_NFIELDS], size_t nc)
{
  if (!ctx || !fields) return IPGPU_ERR_ARG;
  if (!ctx->has_geom) return set_err(ctx, IPGPU_ERR_STATE, "geometry not set");
  if (nc != ghosted_cells(ctx->geom)) return set_err(ctx, IPGPU_ERR_ARG, "ncells_ghosted mismatch");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = ensure_dump(ctx);
  if (rc) return rc;
  double *soa = nullptr;
  HIPCHK(ctx, hipMalloc((void **)&soa, nc * IPGPU_NFIELDS * sizeof(double)));
  for (int f = 0; f < IPGPU_NFIELDS; f++) {
    if (!fields[f]) {
      hipFree(soa);
      return set_err(ctx, IPGPU_ERR_ARG, "null field pointer");
    }
    HIPCHK(ctx, hipMemcpyAsync(soa + f * nc, fields[f], nc * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  k_pack<<<(unsigned)((nc + 255) / 256), 256, 0, ctx->stream>>>(nc, soa, ctx->d_dump);
  ctx->has_units = false; /* the electron model behind these fields is not known */
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipFree(soa));
  return IPGPU_OK;
}

int ipgpu_upload_prims(ipgpu_ctx *ctx, const double *prims, int nprim, const ipgpu_units *u)
{
  if (!ctx || !prims || !u || nprim < 8) return IPGPU_ERR_ARG;
  if (!ctx->has_geom) return set_err(ctx, IPGPU_ERR_STATE, "geometry not set");
  const ipgpu_geom &G = ctx->geom;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc = ensure_dump(ctx);
  if (rc) return rc;
  size_t nc = ghosted_cells(G);
  size_t nin = (size_t)G.N1 * G.N2 * G.N3;
  double *dp = nullptr;
  HIPCHK(ctx, hipMalloc((void **)&dp, nin * nprim * sizeof(double)));
  HIPCHK(ctx, hipMemcpyAsync(dp, prims, nin * nprim * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  /* set_units, model/iharm/model.c:595-608 */
  double MBH = u->MBH_solar * 1.989e33;
  double L_unit = 6.6742e-8 * MBH / (CL * CL);
  double RHO_unit = u->M_unit / pow(L_unit, 3);
  double B_unit = CL * sqrt(4. * M_PI * RHO_unit);
  k_derive_b<<<(unsigned)((nin + 127) / 128), 128, 0, ctx->stream>>>(G, B_unit, nprim, dp, ctx->d_dump);
  int nr = G.N2 * G.N3, np = (G.N1 + 2) * G.N3, nf = (G.N1 + 2) * (G.N2 + 2);
  k_bc_radial<<<(nr + 255) / 256, 256, 0, ctx->stream>>>(G, ctx->d_dump);
  k_bc_polar<<<(np + 255) / 256, 256, 0, ctx->stream>>>(G, ctx->d_dump);
  k_bc_phi<<<(nf + 255) / 256, 256, 0, ctx->stream>>>(G, ctx->d_dump);
  k_derive_thermo<<<(unsigned)((nc + 255) / 256), 256, 0, ctx->stream>>>(nc, *u, RHO_unit, B_unit, ctx->d_dump);
  HIPCHK(ctx, hipGetLastError());
  ctx->units = *u;
  ctx->has_units = true;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipFree(dp));
  return IPGPU_OK;
}

int ipgpu_download_fields(ipgpu_ctx *ctx, double *out11, size_t nc)
{
  if (!ctx || !out11) return IPGPU_ERR_ARG;
  if (!ctx->d_dump || nc != ctx->dump_cells) return set_err(ctx, IPGPU_ERR_STATE, "no dump on device");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  double *soa = nullptr;
  HIPCHK(ctx, hipMalloc((void **)&soa, nc * 11 * sizeof(double)));
  k_unpack<<<(unsigned)((nc + 255) / 256), 256, 0, ctx->stream>>>(nc, ctx->d_dump, soa);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(out11, soa, nc * 11 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipFree(soa));
  return IPGPU_OK;
}

static int run_trace(ipgpu_ctx *ctx, const ipgpu_image &P, int nrays, int cap, bool turns)
{
  size_t need = (size_t)nrays * cap * REC;
  int rc = grow(ctx, ctx->d_traj, ctx->traj_doubles, need);
  if (rc) return rc;
  if (turns) {
    rc = grow(ctx, ctx->d_nturns, ctx->nturns_ints, (size_t)nrays * cap);
    if (rc) return rc;
  }
  HIPCHK(ctx, hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_counts, 0, 2 * sizeof(unsigned long long), ctx->stream));
  if (turns)
    k_trace<<<nrays / WAVE, WAVE, 0, ctx->stream>>>(ctx->geom, P, ctx->pix, nrays, cap, ctx->d_traj, ctx->d_nstep_ray,
                                                     ctx->d_nturns, ctx->d_status, ctx->d_counts);
  else
    k_trace_pair<<<2 * nrays / WAVE, WAVE, 0, ctx->stream>>>(ctx->geom, P, ctx->pix, nrays, cap, ctx->d_traj,
                                                             ctx->d_nstep_ray, ctx->d_status, ctx->d_counts);
  HIPCHK(ctx, hipGetLastError());
  return 0;
}

/* Geodesics (retraced once with room for the longest ray if the record
 * capacity was short), packet offsets and the step packets of one image;
 * with aux, also the adjoint records.  Records events ev[0..2]. */
static int prepare_packets(ipgpu_ctx *ctx, const ipgpu_image &P, int nrays, bool aux, int &cap,
                           unsigned long long &total, int &stride)
{
  int rc;
  bool turns = P.target_nturns >= 0;
  cap = ctx->traj_cap_steps < P.maxnstep ? ctx->traj_cap_steps : P.maxnstep;
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  if ((rc = run_trace(ctx, P, nrays, cap, turns))) return rc;
  int status[NSTATUS];
  total = 0;
  HIPCHK(ctx, hipMemcpyAsync(status, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&total, ctx->d_counts, sizeof(total), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (status[0] > 0) {
    /* some rays outgrew the per-ray record capacity: retrace with room for
     * the longest one (bounded by maxnstep) */
    std::vector<int> ns(nrays);
    HIPCHK(ctx, hipMemcpy(ns.data(), ctx->d_nstep_ray, nrays * sizeof(int), hipMemcpyDeviceToHost));
    int mx = 0;
    for (int v : ns) mx = v > mx ? v : mx;
    cap = mx + 1 < P.maxnstep ? mx + 1 : P.maxnstep;
    ctx->traj_cap_steps = cap;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    if ((rc = run_trace(ctx, P, nrays, cap, turns))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(&total, ctx->d_counts, sizeof(total), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  stride = P.only_unpolarized ? PK_UNPOL : PK;
  if (aux && (rc = grow(ctx, ctx->d_aux, ctx->aux_doubles, (size_t)(total > 0 ? total : 1) * AX))) return rc;
  if ((rc = grow(ctx, ctx->d_pkt, ctx->pkt_doubles, (size_t)(total > 0 ? total : 1) * stride))) return rc;
  if ((rc = grow(ctx, ctx->d_off, ctx->cap_off, (size_t)nrays + 1))) return rc;
  static const int init[NSTATUS] = {0, 0x7fffffff, 0x7fffffff, 0};
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_status, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  k_scan<<<1, SCAN_THREADS, 0, ctx->stream>>>(nrays, ctx->d_nstep_ray, ctx->d_off);
  HIPCHK(ctx, hipGetLastError());
  if (total > 0) {
    unsigned nb = (unsigned)((total + PREP_BLOCK - 1) / PREP_BLOCK);
    if ((rc = grow(ctx, ctx->d_blk, ctx->cap_blk, (size_t)nb))) return rc;
    k_blockmap<<<(nrays + 255) / 256, 256, 0, ctx->stream>>>(nrays, ctx->d_off, ctx->d_blk);
    HIPCHK(ctx, hipGetLastError());
    auto kp = P.transfer_fp32 ? (ctx->gather_stage ? k_prep<true, true> : k_prep<true, false>)
                              : (ctx->gather_stage ? k_prep<false, true> : k_prep<false, false>);
    kp<<<nb, PREP_BLOCK, 0, ctx->stream>>>(ctx->geom, ctx->model, P, nrays, cap, stride, (long long)total,
                                               ctx->d_off, ctx->d_blk, ctx->d_nstep_ray, ctx->d_traj,
                                               turns ? ctx->d_nturns : nullptr, ctx->d_dump, ctx->d_pkt,
                                               aux ? ctx->d_aux : nullptr, ctx->r_d32);
    HIPCHK(ctx, hipGetLastError());
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  return 0;
}

/* Segment tables of the rays of a list whose chains are composed (rays
 * longer than seg_min_steps); off[k]: position of element k's first packet
 * in doubles; base[k] = first segment of element k or -1. */
static void build_segments(ipgpu_ctx::SegSet &S, const int *list, int n, const long long *off,
                           const std::vector<int> &ns, int seg_min, int seg_len, int maxnstep, int *base,
                           int covered = 0)
{
  /* covered: a multiple of seg_len; the segments of each ray's first
   * covered steps from the camera are formed elsewhere and left out (its
   * packets here are those of the other ns - covered steps) */
  S.off.clear();
  S.q0.clear();
  S.q1.clear();
  for (int k = 0; k < n; k++) {
    const int v = ns[list[k]];
    base[k] = -1;
    if (v > seg_min && v < maxnstep - 1) {
      base[k] = (int)S.off.size();
      const int nseg = (v - covered + seg_len - 1) / seg_len;
      for (int sg = 0; sg < nseg; sg++) {
        int q0, q1;
        segment_bounds(v, seg_len, sg, q0, q1);
        S.off.push_back(off[k]);
        S.q0.push_back(q0);
        S.q1.push_back(q1);
      }
    }
  }
}

static int upload_segments(ipgpu_ctx *ctx, ipgpu_ctx::SegSet &S, hipStream_t st)
{
  const size_t nseg = S.off.size();
  if (nseg == 0) return 0;
  int rc;
  if ((rc = grow(ctx, S.d_off, S.cap_off, nseg)) || (rc = grow(ctx, S.d_q0, S.cap_q0, nseg))
      || (rc = grow(ctx, S.d_q1, S.cap_q1, nseg)) || (rc = grow(ctx, S.d_flag, S.cap_flag, nseg))
      || (rc = grow(ctx, S.d_map, S.cap_map, nseg * SEGCOL * NPAR)))
    return rc;
  HIPCHK(ctx, hipMemcpyAsync(S.d_off, S.off.data(), nseg * sizeof(long long), hipMemcpyHostToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(S.d_q0, S.q0.data(), nseg * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(S.d_q1, S.q1.data(), nseg * sizeof(int), hipMemcpyHostToDevice, st));
  return 0;
}

/* Segment maps of the rays' first steps formed ahead (k_prep_prefix and
 * k_seg on another stream): element k's npre maps at map + idx[k] npre,
 * complete when ev has fired. */
struct PreMaps {
  const int *idx;
  const double *map;
  const int *flag;
  int npre;
  hipEvent_t ev;
};

/* Step packets, segment maps and chains of the n rays of a list (device
 * list dl sorted longest first, wave rows W and packet-kernel tiles TT with
 * nw + 1 entries each, rows and tiles in all), on stream st; packets
 * wave-interleaved at pkt, block map at blk.  pre (compact lists only):
 * the maps of the rays' first steps, formed ahead. */
static int transfer_list(ipgpu_ctx *ctx, const ipgpu_image &P, hipStream_t st, int n, const int *dl,
                         const long long *W, const long long *TT, long long tiles, int nw, long long rows, int *blk,
                         double *pkt, const int *dbase, ipgpu_ctx::SegSet &S, int cap, int stride, bool compact,
                         const PreMaps *pre = nullptr, hipEvent_t after_prep = nullptr,
                         const long long *rbase = nullptr, const double *trajO = nullptr)
{
  if (n <= 0) return 0;
  if (compact) {
    /* every chain composed: W holds the n + 1 packet offsets, rows the total */
    if (rows > 0) {
      const unsigned nb = (unsigned)((rows + PREP_BLOCK - 1) / PREP_BLOCK);
      k_blockmap_list<<<(n + 255) / 256, 256, 0, st>>>(n, W, blk);
      HIPCHK(ctx, hipGetLastError());
      auto kp = P.transfer_fp32 ? (ctx->gather_stage ? k_prep_compact<true, true> : k_prep_compact<true, false>)
                                : (ctx->gather_stage ? k_prep_compact<false, true> : k_prep_compact<false, false>);
      kp<<<nb, PREP_BLOCK, 0, st>>>(ctx->geom, ctx->model, P, dl, n, W, rows, blk, cap, ctx->d_nstep_ray,
                                    ctx->d_traj, rbase, trajO, ctx->d_dump, pkt, ctx->r_d32);
      HIPCHK(ctx, hipGetLastError());
    }
    const int nseg = (int)S.off.size();
    if (nseg > 0) {
      long long lanes = (long long)nseg * SEGCOL;
      k_seg<1><<<(unsigned)((lanes + WAVE - 1) / WAVE), WAVE, 0, st>>>(nseg, S.d_off, S.d_q0, S.d_q1,
                                                                      pkt, S.d_map, S.d_flag, 1);
      HIPCHK(ctx, hipGetLastError());
    }
    if (pre) HIPCHK(ctx, hipStreamWaitEvent(st, pre->ev, 0));
    k_chain_list<true><<<(n + WAVE - 1) / WAVE, WAVE, 0, st>>>(ctx->geom, ctx->model, P, ctx->pix, dl, n, W, cap, stride,
                                                              ctx->d_nstep_ray, ctx->d_traj, rbase, trajO, ctx->d_dump,
                                                              pkt, dbase, ctx->seg_len, S.d_map, S.d_flag, ctx->d_out7,
                                                              ctx->d_flag_ray, ctx->d_status, ctx->d_counts,
                                                              pre ? pre->idx : nullptr, pre ? pre->map : nullptr,
                                                              pre ? pre->flag : nullptr, pre ? pre->npre : 0);
    HIPCHK(ctx, hipGetLastError());
    return 0;
  }
  if (rows > 0) {
    const unsigned nb = (unsigned)((tiles + TPB - 1) / TPB);   /* blk holds nb entries */
    k_blockmap_tiles<<<(nw + 255) / 256, 256, 0, st>>>(nw, TT, blk);
    HIPCHK(ctx, hipGetLastError());
    auto kp = P.transfer_fp32 ? (ctx->gather_stage ? k_prep_list<true, true> : k_prep_list<true, false>)
                              : (ctx->gather_stage ? k_prep_list<false, true> : k_prep_list<false, false>);
    kp<<<nb, PREP_BLOCK, 0, st>>>(ctx->geom, ctx->model, P, dl, n, W, TT, nw, tiles, blk, cap, stride,
                                  ctx->d_nstep_ray, ctx->d_traj, rbase, trajO, ctx->d_dump, pkt, ctx->r_d32);
    HIPCHK(ctx, hipGetLastError());
  }
  if (after_prep) HIPCHK(ctx, hipEventRecord(after_prep, st));
  const int nseg = (int)S.off.size();
  if (nseg > 0) {
    long long lanes = (long long)nseg * SEGCOL;
    k_seg<ILV><<<(unsigned)((lanes + WAVE - 1) / WAVE), WAVE, 0, st>>>(nseg, S.d_off, S.d_q0, S.d_q1,
                                                                      pkt, S.d_map, S.d_flag, 1);
    HIPCHK(ctx, hipGetLastError());
  }
  /* composed rays (the longest, at the front of the list) and the rest */
  if (nseg > 0) {
    k_chain_list<true><<<(n + WAVE - 1) / WAVE, WAVE, 0, st>>>(ctx->geom, ctx->model, P, ctx->pix, dl, n, W, cap, stride,
                                                              ctx->d_nstep_ray, ctx->d_traj, rbase, trajO, ctx->d_dump,
                                                              pkt, dbase, ctx->seg_len, S.d_map, S.d_flag, ctx->d_out7,
                                                              ctx->d_flag_ray, ctx->d_status, ctx->d_counts, nullptr,
                                                              nullptr, nullptr, 0);
    HIPCHK(ctx, hipGetLastError());
  }
  k_chain_list<false><<<(n + WAVE - 1) / WAVE, WAVE, 0, st>>>(ctx->geom, ctx->model, P, ctx->pix, dl, n, W, cap, stride,
                                                             ctx->d_nstep_ray, ctx->d_traj, rbase, trajO, ctx->d_dump,
                                                             pkt, nseg > 0 ? dbase : nullptr, ctx->seg_len, S.d_map,
                                                             S.d_flag, ctx->d_out7, ctx->d_flag_ray, ctx->d_status,
                                                             ctx->d_counts, nullptr, nullptr, nullptr, 0);
  HIPCHK(ctx, hipGetLastError());
  return 0;
}

/* Step-count bands of the streamed render's lists (2^LIST_BAND steps). */
constexpr int LIST_BAND = 6;

/* Sorts list[0..n) by ns >> band, longest first (counting sort; ties in
 * list order).  With band > 0 the rays of a wave of 64 are neighbours in the
 * list's original order (ray index) within a band of similar lengths: the
 * packet kernel's tiles of 8 rays then follow nearby trajectories through
 * the same dump cells (on the 160x160 image, 64-step bands cut k_prep_list's
 * L1 misses by a third and its time by 2%, against the exact order). */
static void sort_longest_first(int *list, int n, const std::vector<int> &ns, std::vector<int> &cnt,
                               std::vector<int> &tmp, int band = 0)
{
  int mx = 0;
  for (int k = 0; k < n; k++) mx = (ns[list[k]] >> band) > mx ? (ns[list[k]] >> band) : mx;
  cnt.assign((size_t)mx + 2, 0);
  for (int k = 0; k < n; k++) cnt[mx - (ns[list[k]] >> band) + 1]++;
  for (int v = 1; v <= mx + 1; v++) cnt[v] += cnt[v - 1];
  tmp.resize(n);
  for (int k = 0; k < n; k++) tmp[cnt[mx - (ns[list[k]] >> band)]++] = list[k];
  for (int k = 0; k < n; k++) list[k] = tmp[k];
}

/* Full-length record regions for the rays the unbounded phase of a
 * streamed render traces (0: no pool; the per-ray slots then follow the
 * longest ray, as the one-pass render's do): the last render's share of
 * such rays with a quarter more, at most a quarter of the render memory. */
static int pool_regions(const ipgpu_ctx *ctx, const ipgpu_image &P, int nrays)
{
  if (!ctx->pool_on || ctx->pool_max == 0 || P.target_nturns >= 0 || ctx->trace_budget <= 0
      || ctx->trace_budget + 1 >= P.maxnstep)
    return 0;
  const double reg_bytes = (double)(P.maxnstep + 1) * REC * sizeof(double);
  const double want = (double)nrays * ctx->live_frac * 1.25 + 64.;
  const double lim = 0.25 * (double)ctx->render_bytes / reg_bytes;
  double n = want < lim ? want : lim;
  if (n > nrays) n = nrays;
  if (ctx->pool_max > 0 && n > ctx->pool_max) n = ctx->pool_max;
  return n >= 1. ? (int)n : 0;
}

/* Streamed render of one image.  The geodesics are integrated in phases
 * with growing step budgets (trace_budget/4, /2, trace_budget, then to the
 * end), all queued at once on the main stream; each phase ends with a copy
 * of the step counts and live flags to pinned host memory.  As soon as a
 * bounded phase has ended, the rays it finished are prepared and chained on
 * a stream of their own, concurrent with the later phases.  The rays alive
 * after the last bounded phase (the longest) are prepared, segmented and
 * chained on the main stream after their trace; the streams join before
 * k_finalize.  The per-ray arithmetic is that of the one-pass render (a
 * resumed trace continues from its stored record, which holds X and K
 * exactly).  A ray that outgrows the record capacity (cap records per ray)
 * is continued from its last stored record in a full-length region of its
 * own before its transfer. */
static int render_streamed(ipgpu_ctx *ctx, const ipgpu_image &P, int nrays, int cap, int nreg)
{
  int rc;
  const int stride = P.only_unpolarized ? PK_UNPOL : PK;
  const bool segm = !P.only_unpolarized && !P.stokes_floors;
  const int seg_min = segm ? ctx->seg_min_steps : 0x7fffffff;
  hipStream_t sA = ctx->stream;
  hipEvent_t *ev = ctx->ev;
  /* budgets of the bounded phases: trace_budget j / nb for j = 1 .. nb
   * (nb = trace_phases), those below 64 steps dropped */
  int bud[MAXPH - 1], K = 0;
  const int nb = ctx->trace_phases < 1 ? 1 : (ctx->trace_phases > MAXPH - 1 ? MAXPH - 1 : ctx->trace_phases);
  for (int j = 1; j <= nb; j++) {
    int b = (int)((long long)ctx->trace_budget * j / nb);
    if (b >= 64 && (K == 0 || b > bud[K - 1])) bud[K++] = b;
  }
  if (K == 0) bud[K++] = ctx->trace_budget;
  /* events: trace kernels [2k, 2k+1] (ev[0] opens phase 0), end of the
   * finished-ray work of phase k 15+k, first finished-ray work 19, longest
   * rays' work done 20, join 21, finalize 22 */
  if ((rc = grow_hr(ctx, ctx->d_traj, ctx->traj_doubles, (size_t)nrays * cap * REC))) return rc;
  if ((rc = grow(ctx, ctx->d_live, ctx->cap_live, (size_t)nrays))) return rc;
  if ((rc = grow(ctx, ctx->d_list, ctx->cap_list, (size_t)nrays))) return rc;
  if ((rc = grow(ctx, ctx->d_offl, ctx->cap_offl, 2 * ((size_t)nrays + MAXPH + 1)))) return rc;
  if ((rc = grow(ctx, ctx->d_segbase, ctx->cap_segbase, (size_t)nrays))) return rc;
  if (ctx->last_total > 0) {
    if ((rc = grow_hr(ctx, ctx->d_pkt, ctx->pkt_doubles, (size_t)ctx->last_total * stride))) return rc;
    /* tiles: rows / 4 + 2 per list wave (ceil to 8 steps) per block of 4 */
    if ((rc = grow(ctx, ctx->d_blk, ctx->cap_blk,
                   (size_t)(ctx->last_total / PREP_BLOCK + 2 * ((long long)nrays / ILV + 2 * MAXPH)))))
      return rc;
  }
  /* pinned, device-mapped, coherent host arena: the trace kernels write
   * each phase's step counts and live flags into it (a copy command would
   * hold the host until the phase ends) */
  const size_t nI = (size_t)nrays;
  const size_t need = sizeof(int) * nI * (2 * (MAXPH) + 4) + 2 * sizeof(long long) * (nI + MAXPH + 1)
      + sizeof(long long) * nI;
  if (ctx->h_pin_bytes < need) {
    if (ctx->h_pin) hipHostFree(ctx->h_pin);
    ctx->h_pin = nullptr;
    ctx->h_pin_bytes = 0;
    HIPCHK(ctx, hipHostMalloc(&ctx->h_pin, need, hipHostMallocMapped | hipHostMallocCoherent));
    ctx->h_pin_bytes = need;
  }
  char *d_pin = nullptr;
  HIPCHK(ctx, hipHostGetDevicePointer((void **)&d_pin, ctx->h_pin, 0));
  long long *h_off = (long long *)ctx->h_pin;
  const size_t ns_bytes = 2 * sizeof(long long) * (nI + MAXPH + 1);   /* W, then TT */
  int *h_ns = (int *)((char *)ctx->h_pin + ns_bytes); /* [K+1][nrays] */
  int *h_live = h_ns + (size_t)MAXPH * nI;            /* [K][nrays] */
  int *h_list = h_live + (size_t)MAXPH * nI;
  int *h_base = h_list + nI;
  int *h_plist = h_base + nI, *h_pidx = h_plist + nI;
  int *dh_ns = (int *)(d_pin + ns_bytes), *dh_live = dh_ns + (size_t)MAXPH * nI;
  /* pooled regions' offsets per ray, after the int rows (8-byte aligned) */
  const size_t rb_bytes = ns_bytes + sizeof(int) * nI * (2 * (MAXPH) + 4);
  const long long *h_rbase = (const long long *)((char *)ctx->h_pin + rb_bytes);
  long long *dh_rbase = (long long *)(d_pin + rb_bytes);
  /* the record pool of the unbounded phase */
  const long long reg = (long long)(P.maxnstep + 1) * REC;   /* doubles per region */
  const long long pool_doubles = (long long)nreg * reg;
  if (nreg > 0) {
    if ((rc = grow_hr(ctx, ctx->d_trajO, ctx->trajO_doubles, (size_t)pool_doubles))
        || (rc = grow(ctx, ctx->d_pool_list, ctx->cap_pool_list, (size_t)nreg))
        || (rc = grow(ctx, ctx->d_pool_ctr, ctx->cap_pool_ctr, (size_t)1))
        || (rc = grow(ctx, ctx->d_rbase, ctx->cap_rbase, nI)))
      return rc;
  }
  ctx->last_pool_regions = nreg;
  /* the first Ppre steps (a multiple of the segment length, records all
   * stored by the bounded phases) of the rays the unbounded phase traces
   * are prepared and composed on stream streamF[K] while it runs */
  const int Ppre = K >= 1 ? ((bud[K - 1] - 1) / ctx->seg_len) * ctx->seg_len : 0;
  const bool pre_on = ctx->tail_prefix && segm && K >= 1 && K <= MAXPH - 2 && seg_min <= bud[K - 1]
      && Ppre >= ctx->seg_len && Ppre + 1 < cap;
  bool pre_launched = false, prep_ev = false;
  int pre_npre = 0;

  static const int init[NSTATUS] = {0, 0x7fffffff, 0x7fffffff, 0};
  HIPCHK(ctx, hipEventRecord(ev[0], sA));
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_status, init, sizeof(init), hipMemcpyHostToDevice, sA));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_counts, 0, 2 * sizeof(unsigned long long), sA));
  /* all trace phases, queued at once; phase k writes the step counts (and
   * live flags) of the rays it integrated to the host arena's row k (the
   * live-flag row of the last, unbounded phase is scratch) */
  if (nreg > 0) HIPCHK(ctx, hipMemsetAsync(ctx->d_pool_ctr, 0, sizeof(unsigned int), sA));
  for (int k = 0; k <= K; k++) {
    const bool pooled = k == K && nreg > 0;
    if (pooled) {
      /* the rays still live get their regions, and their records so far */
      k_pool_assign<<<(nrays + 255) / 256, 256, 0, sA>>>(nrays, ctx->d_live, nreg, reg, ctx->d_pool_ctr,
                                                        ctx->d_rbase, dh_rbase, ctx->d_pool_list);
      HIPCHK(ctx, hipGetLastError());
      k_pool_copy<<<nreg, 256, 0, sA>>>(nreg, ctx->d_pool_ctr, ctx->d_pool_list, ctx->d_nstep_ray, cap,
                                          ctx->d_traj, ctx->d_trajO, reg);
      HIPCHK(ctx, hipGetLastError());
    }
    if (k > 0) HIPCHK(ctx, hipEventRecord(ev[2 * k], sA));
    int *hn = dh_ns + (size_t)k * nI, *hl = dh_live + (size_t)(k < K ? k : MAXPH - 1) * nI;
    if (k == 0)
      k_trace_budget<<<2 * nrays / WAVE, WAVE, 0, sA>>>(ctx->geom, P, ctx->pix, nrays, cap, bud[0], ctx->d_traj,
                                                         ctx->d_nstep_ray, ctx->d_live, hn, hl, ctx->d_counts);
    else
      k_trace_resume_mask<<<2 * nrays / WAVE, WAVE, 0, sA>>>(
          ctx->geom, P, nrays, cap, k < K ? bud[k] : -1, ctx->d_traj, ctx->d_nstep_ray, ctx->d_live, hn, hl,
          ctx->d_counts, pooled ? ctx->d_rbase : nullptr, pooled ? ctx->d_trajO : nullptr, P.maxnstep + 1);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ev[2 * k + 1], sA));
  }

  /* the rays of each phase as it ends: list region [start, start + n),
   * offsets at start + k, packets bump-allocated in d_pkt (phases up to K-1)
   * or d_pkt2 (the last) */
  std::vector<int> nsv(nI);
  int start = 0;
  long long pofs = 0, bofs = 0, total = 0, total_f = 0, pre_slots = 0;
  bool any_f = false;
  for (int k = 0; k <= K; k++) {
    HIPCHK(ctx, hipEventSynchronize(ev[2 * k + 1]));
    const volatile int *nsk = h_ns + (size_t)k * nI;
    int n = 0;
    for (int r = 0; r < nrays; r++) {
      /* rays integrated in phase k (all in phase 0, else those live after
       * phase k - 1) that it finished */
      if (k > 0 && !((volatile int *)h_live)[(size_t)(k - 1) * nI + r]) continue;
      if (k < K && ((volatile int *)h_live)[(size_t)k * nI + r]) continue;
      if (start + n >= nrays) return set_err(ctx, IPGPU_ERR_STATE, "streamed render: inconsistent phase lists");
      h_list[start + n++] = r;
      nsv[r] = nsk[r];
    }
    const long long *rbase = nullptr;
    if (k == K) {
      /* rays past the record capacity without a pooled region: their first
       * cap records move to full-length regions (after the pool's), where
       * their traces continue */
      std::vector<int> &ov = ctx->h_ovl;
      std::vector<long long> &oo = ctx->h_ovoff;
      ov.clear();
      oo.assign(1, pool_doubles / REC);
      int mx = 0;
      for (int q = 0; q < n; q++) {
        const int r = h_list[start + q], v = nsv[r];
        mx = v > mx ? v : mx;
        const bool in_pool = nreg > 0 && ((const volatile long long *)h_rbase)[r] >= 0;
        if (v >= cap && v < P.maxnstep - 1 && !in_pool) {
          ov.push_back(r);
          oo.push_back(oo.back() + v + 1);
        }
      }
      ctx->last_max_steps = mx;
      if (nreg > 0) {
        /* the share of rays the unbounded phase traced sizes the next pool */
        ctx->live_frac = std::max((double)n / nrays, 1.0 / 4096);
        ctx->last_pool_used = std::min(n, nreg);
        rbase = ctx->d_rbase;
      }
      if (!ov.empty()) {
        const int no = (int)ov.size();
        if ((rc = grow_keep(ctx, ctx->d_trajO, ctx->trajO_doubles, (size_t)oo.back() * REC, (size_t)pool_doubles,
                            sA))
            || (rc = grow(ctx, ctx->d_ovl, ctx->cap_ovl, (size_t)no))
            || (rc = grow(ctx, ctx->d_ovoff, ctx->cap_ovoff, (size_t)no))
            || (rc = grow(ctx, ctx->d_rbase, ctx->cap_rbase, nI)))
          return rc;
        if (nreg == 0) HIPCHK(ctx, hipMemsetAsync(ctx->d_rbase, 0xff, nI * sizeof(long long), sA));
        HIPCHK(ctx, hipMemsetAsync(ctx->d_status, 0, sizeof(int), sA));
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_ovl, ov.data(), no * sizeof(int), hipMemcpyHostToDevice, sA));
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_ovoff, oo.data(), no * sizeof(long long), hipMemcpyHostToDevice, sA));
        const long long cells = (long long)no * cap * REC;
        k_slot_copy<<<(unsigned)((cells + 255) / 256 < 65536 ? (cells + 255) / 256 : 65536), 256, 0, sA>>>(
            ctx->d_ovl, no, ctx->d_ovoff, cap, ctx->d_traj, ctx->d_trajO);
        HIPCHK(ctx, hipGetLastError());
        k_trace_continue<<<(2 * no + WAVE - 1) / WAVE, WAVE, 0, sA>>>(ctx->geom, P, ctx->d_ovl, no, ctx->d_ovoff, cap,
                                                                      ctx->d_nstep_ray, ctx->d_trajO, ctx->d_rbase,
                                                                      ctx->d_status);
        HIPCHK(ctx, hipGetLastError());
        rbase = ctx->d_rbase;
      }
      ctx->last_overflow = (int)ov.size();
    }
    /* longest first.  Lists whose chains are all composed from segments
     * (the last phase's, when the segment threshold is at most the last
     * budget) keep each ray's packets contiguous (offsets in W); the others
     * are cut into waves of 64 with wave-interleaved packets (rows per wave
     * in W). */
    bool compact = (k == K) && segm && n > 0;
    for (int q = 0; compact && q < n; q++) compact = nsv[h_list[start + q]] > seg_min;
    sort_longest_first(h_list + start, n, nsv, ctx->h_cnt, ctx->h_tmp, compact ? 0 : LIST_BAND);
    /* the rays' first Ppre steps were composed ahead: their packets here are
     * those of the other steps */
    const bool use_pre = pre_launched && compact;
    const int covered = use_pre ? Ppre : 0;
    if (use_pre)
      for (int q = 0; q < n; q++) {
        h_pidx[q] = ctx->h_ppos[h_list[start + q]];
        if (h_pidx[q] < 0) return set_err(ctx, IPGPU_ERR_STATE, "streamed render: inconsistent prefix list");
      }
    const int nw = compact ? n : (n + ILV - 1) / ILV;
    long long *W = h_off + start + k;
    W[0] = 0;
    if (compact)
      for (int q = 0; q < n; q++) W[q + 1] = W[q] + nsv[h_list[start + q]] - covered;
    else
      for (int c = 0; c < nw; c++) {
        /* the wave's rows: its longest ray's steps */
        int mx = 0;
        for (int q = c * ILV; q < n && q < (c + 1) * ILV; q++) mx = std::max(mx, nsv[h_list[start + q]]);
        W[c + 1] = W[c] + mx;
      }
    const long long rows = W[nw];
    const long long T = compact ? rows : rows * ILV;   /* packet slots, padding included */
    long long *TT = W + (nI + MAXPH + 1);
    TT[0] = 0;
    if (!compact)
      for (int c = 0; c < nw; c++) TT[c + 1] = TT[c] + (long long)(ILV / TL) * ((W[c + 1] - W[c] + TQ - 1) / TQ);
    const long long tiles = compact ? 0 : TT[nw];
    total += T;
    if (k < K) total_f += T;
    /* first packet of each element, in doubles */
    std::vector<long long> &eoff = ctx->h_eoff;
    eoff.resize(n);
    for (int q = 0; q < n; q++) eoff[q] = compact ? W[q] * PK : W[q / ILV] * stride * ILV + (long long)pk_lane<ILV>(q % ILV);
    build_segments(ctx->segset[k], h_list + start, n, eoff.data(), nsv, seg_min, ctx->seg_len, P.maxnstep,
                   h_base + start, covered);
    hipStream_t st = (k < K) ? ctx->streamF[k] : sA;
    int *dl = ctx->d_list + start, *dbase = ctx->d_segbase + start;
    long long *dW = ctx->d_offl + start + k, *dTT = dW + (nI + MAXPH + 1);
    double *pkt;
    int *blk;
    if (n > 0) {
      const size_t nb = compact ? (size_t)((T + PREP_BLOCK - 1) / PREP_BLOCK) : (size_t)((tiles + TPB - 1) / TPB);
      if (k < K) {
        if ((size_t)(pofs + T) * stride > ctx->pkt_doubles || bofs + nb > ctx->cap_blk) {
          /* the arenas are full: earlier phases finish first */
          for (int q = 0; q < k; q++) HIPCHK(ctx, hipStreamSynchronize(ctx->streamF[q]));
          pofs = bofs = 0;
          const size_t want = (size_t)(T > ctx->last_total ? T : ctx->last_total);
          if ((rc = grow_hr(ctx, ctx->d_pkt, ctx->pkt_doubles, (want > 0 ? want : 1) * stride))) return rc;
          if ((rc = grow(ctx, ctx->d_blk, ctx->cap_blk, want / PREP_BLOCK + 2 * ((size_t)nrays / ILV + 2 * MAXPH))))
            return rc;
        }
        pkt = ctx->d_pkt + (size_t)pofs * stride;
        blk = ctx->d_blk + bofs;
        pofs += T;
        bofs += nb;
        if (!any_f) HIPCHK(ctx, hipEventRecord(ev[19], st));
        any_f = true;
      } else {
        if ((rc = grow_hr(ctx, ctx->d_pkt2, ctx->pkt2_doubles, (size_t)(T > 0 ? T : 1) * stride))) return rc;
        if ((rc = grow(ctx, ctx->d_blk2, ctx->cap_blk2, nb > 0 ? nb : 1))) return rc;
        pkt = ctx->d_pkt2;
        blk = ctx->d_blk2;
      }
      HIPCHK(ctx, hipMemcpyAsync(dl, h_list + start, n * sizeof(int), hipMemcpyHostToDevice, st));
      HIPCHK(ctx, hipMemcpyAsync(dW, W, (nw + 1) * sizeof(long long), hipMemcpyHostToDevice, st));
      if (!compact) HIPCHK(ctx, hipMemcpyAsync(dTT, TT, (nw + 1) * sizeof(long long), hipMemcpyHostToDevice, st));
      if (!ctx->segset[k].off.empty())
        HIPCHK(ctx, hipMemcpyAsync(dbase, h_base + start, n * sizeof(int), hipMemcpyHostToDevice, st));
      if ((rc = upload_segments(ctx, ctx->segset[k], st))) return rc;
      PreMaps pm{ctx->d_pidx, ctx->segpre.d_map, ctx->segpre.d_flag, pre_npre, ev[23]};
      if (use_pre) {
        if ((rc = grow(ctx, ctx->d_pidx, ctx->cap_pidx, (size_t)n))) return rc;
        pm.idx = ctx->d_pidx;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_pidx, h_pidx, n * sizeof(int), hipMemcpyHostToDevice, st));
      }
      if ((rc = transfer_list(ctx, P, st, n, dl, dW, dTT, tiles, nw, rows, blk, pkt, dbase, ctx->segset[k], cap,
                              stride, compact, use_pre ? &pm : nullptr,
                              (k == K - 1 && pre_on && !compact) ? ev[10] : nullptr, rbase, ctx->d_trajO)))
        return rc;
      prep_ev = k == K - 1 && pre_on && !compact;
    }
    if (k < K) HIPCHK(ctx, hipEventRecord(ev[15 + k], st));
    if (k == K - 1 && pre_on) {
      /* the rays still traced, whose first Ppre steps are all stored */
      int np = 0;
      ctx->h_ppos.assign(nI, -1);
      for (int r = 0; r < nrays; r++)
        if (((volatile int *)h_live)[(size_t)(K - 1) * nI + r]) {
          ctx->h_ppos[r] = np;
          h_plist[np++] = r;
        }
      if (np > 0) {
        const int np1 = Ppre + 1, npre = Ppre / ctx->seg_len;
        ipgpu_ctx::SegSet &S = ctx->segpre;
        S.off.clear();
        S.q0.clear();
        S.q1.clear();
        for (int j = 0; j < np; j++)
          for (int sg = 0; sg < npre; sg++) {
            S.off.push_back((long long)j * np1 * PK);
            S.q0.push_back(1 + sg * ctx->seg_len);
            S.q1.push_back(1 + (sg + 1) * ctx->seg_len);
          }
        hipStream_t sp = ctx->streamF[K];
        /* after the packets of the rays this phase finished, which the
         * chain that ends last waits for */
        if (prep_ev) HIPCHK(ctx, hipStreamWaitEvent(sp, ev[10], 0));
        pre_slots = (long long)np * np1;
        if ((rc = grow(ctx, ctx->d_ppkt, ctx->ppkt_doubles, (size_t)np * np1 * PK))
            || (rc = grow(ctx, ctx->d_plist, ctx->cap_plist, (size_t)np)))
          return rc;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_plist, h_plist, np * sizeof(int), hipMemcpyHostToDevice, sp));
        if ((rc = upload_segments(ctx, S, sp))) return rc;
        const long long slots = (long long)np * np1;
        auto kp = P.transfer_fp32 ? (ctx->gather_stage ? k_prep_prefix<true, true> : k_prep_prefix<true, false>)
                                  : (ctx->gather_stage ? k_prep_prefix<false, true> : k_prep_prefix<false, false>);
        kp<<<(unsigned)((slots + PREP_BLOCK - 1) / PREP_BLOCK), PREP_BLOCK, 0, sp>>>(
            ctx->geom, ctx->model, P, ctx->d_plist, np, np1, cap, ctx->d_traj, ctx->d_dump, ctx->d_ppkt, ctx->r_d32);
        HIPCHK(ctx, hipGetLastError());
        const long long lanes = (long long)S.off.size() * SEGCOL;
        k_seg<1><<<(unsigned)((lanes + WAVE - 1) / WAVE), WAVE, 0, sp>>>((int)S.off.size(), S.d_off, S.d_q0, S.d_q1,
                                                                        ctx->d_ppkt, S.d_map, S.d_flag, 0);
        HIPCHK(ctx, hipGetLastError());
        HIPCHK(ctx, hipEventRecord(ev[23], sp));
        pre_launched = true;
        pre_npre = npre;
      }
    }
    start += n;
  }
  if (!any_f) HIPCHK(ctx, hipEventRecord(ev[19], sA));
  /* the arena of the bounded phases' lists is sized by their packets in the
   * next render; the chunk size by all packets of this one */
  ctx->last_total = total_f;
  ctx->pkt_bytes_per_ray = (double)(total + pre_slots) * stride * sizeof(double) / nrays;
  HIPCHK(ctx, hipEventRecord(ev[20], sA));
  for (int k = 0; k < K; k++) HIPCHK(ctx, hipStreamWaitEvent(sA, ev[15 + k], 0));
  if (pre_launched) HIPCHK(ctx, hipStreamWaitEvent(sA, ev[23], 0));
  HIPCHK(ctx, hipEventRecord(ev[21], sA));
  ctx->nphase = K;
  return 0;
}

static int render_rays(ipgpu_ctx *ctx, const ipgpu_image &P, int nrays);

/* Rays per chunk of a single-image render whose ray records and step
 * packets fit the render memory. */
static long long chunk_rays(const ipgpu_ctx *ctx, const ipgpu_image &P)
{
  int cap = ctx->traj_cap_steps < P.maxnstep ? ctx->traj_cap_steps : P.maxnstep;
  /* with the record pool: slots of the bounded phases' records, plus the
   * pool's share of full-length regions per ray */
  double pool_per_ray = 0.;
  if (pool_regions(ctx, P, 1 << 20) > 0) {
    cap = ctx->trace_budget + 1;
    pool_per_ray = ctx->live_frac * 1.25 * (double)(P.maxnstep + 1) * REC * sizeof(double);
  }
  const int stride = P.only_unpolarized ? PK_UNPOL : PK;
  /* per ray: cap records, packets of the mean step count (with the padding
   * of the interleaved layout and the prefix packets of the longest rays),
   * and the per-ray tables; 85% of the render memory, the rest for the
   * buffers of the last phase's list and the segment maps */
  const double pkt = ctx->pkt_bytes_per_ray > 0. ? 1.1 * ctx->pkt_bytes_per_ray
                                                 : ctx->avg_steps_hint * stride * sizeof(double) * 1.25;
  const double per_ray = (double)cap * REC * sizeof(double) + pool_per_ray + pkt + 512.;
  /* the render memory, or what the device can give the render's arenas
   * (free memory plus what they hold now, less 4 GiB), whichever is smaller */
  double budget = (double)ctx->render_bytes;
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) == hipSuccess) {
    const double held = (double)(ctx->traj_doubles + ctx->pkt_doubles + ctx->pkt2_doubles + ctx->ppkt_doubles
                                 + ctx->trajO_doubles) * sizeof(double);
    const double avail = (double)freeb + held - (double)((size_t)4 << 30);
    if (avail < budget) budget = avail > 0. ? avail : 0.;
  }
  const long long chunk = (long long)(0.85 * budget / per_ray) / WAVE * WAVE;
  return chunk < WAVE ? WAVE : chunk;
}

/* The rays of the image in chunks (whole waves of its 8x8 tiles, as pixel
 * lists) whose ray records and step packets fit the render memory, or all
 * at once when they fit.  A pixel's result does not depend on the other
 * rays of its render, so the image is the same either way. */
static int render_image(ipgpu_ctx *ctx, const ipgpu_image &P, int nrays)
{
  if (chunk_rays(ctx, P) >= nrays) {
    ctx->pix = nullptr;
    ctx->last_chunks = 1;
    int rc = render_rays(ctx, P, nrays);
    if (nrays > 0 && ctx->counts[0] > 0) ctx->avg_steps_hint = (double)ctx->counts[0] / nrays;
    return rc;
  }
  /* tiles in the order q -> q s mod T (s coprime to T, near the golden
   * section of T): any run of consecutive tiles is spread over the image,
   * so every chunk has about the image's mix of short and long rays and the
   * next chunk can be sized from the last */
  std::vector<int> &pl = ctx->h_pixlist;
  pl.resize(nrays);
  const int tiles_y = (P.ny + TILE - 1) / TILE;
  const long long T = nrays / WAVE;
  long long stp = (long long)(0.6180339887498949 * (double)T + 0.5);
  auto gcd = [](long long a, long long b) {
    while (b) {
      long long t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  if (stp < 1) stp = 1;
  while (gcd(stp, T) != 1) stp++;
  for (long long q = 0; q < T; q++) {
    const long long t = (q * stp) % T;
    for (int l = 0; l < WAVE; l++) {
      const int i = (int)(t / tiles_y) * TILE + l / TILE, j = (int)(t % tiles_y) * TILE + l % TILE;
      pl[q * WAVE + l] = (i < P.nx && j < P.ny) ? ((i << 16) | j) : -1;
    }
  }
  int rc;
  if ((rc = grow(ctx, ctx->d_pixlist, ctx->cap_pixlist, (size_t)nrays))) return rc;
  unsigned long long counts[2] = {0, 0};
  float ms_geo = 0, ms_xfer = 0, ms_total = 0;
  int nchunks = 0;
  for (long long r0 = 0; r0 < nrays;) {
    /* each chunk sized with the record capacity and mean step count the
     * previous one left */
    const long long chunk = chunk_rays(ctx, P);
    const int n = (int)(nrays - r0 < chunk ? nrays - r0 : chunk);
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_pixlist + r0, pl.data() + r0, (size_t)n * sizeof(int),
                               hipMemcpyHostToDevice, ctx->stream));
    ctx->pix = ctx->d_pixlist + r0;
    rc = render_rays(ctx, P, n);
    ctx->pix = nullptr;
    nchunks++;
    counts[0] += ctx->counts[0];
    counts[1] += ctx->counts[1];
    ms_geo += ctx->ms_geo;
    ms_xfer += ctx->ms_xfer;
    ms_total += ctx->ms_total;
    if (rc) break;
    if (ctx->counts[0] > 0) ctx->avg_steps_hint = (double)ctx->counts[0] / n;
    r0 += n;
  }
  ctx->counts[0] = counts[0];
  ctx->counts[1] = counts[1];
  ctx->ms_geo = ms_geo;
  ctx->ms_xfer = ms_xfer;
  ctx->ms_total = ms_total;
  ctx->last_chunks = nchunks;
  return rc;
}

int ipgpu_render(ipgpu_ctx *ctx, const ipgpu_image *img, double *image, double *imageS,
                 double *taus, int *nstep, int *flag)
{
  if (!ctx || !img) return IPGPU_ERR_ARG;
  if (!ctx->has_geom || !ctx->has_model) return set_err(ctx, IPGPU_ERR_STATE, "geometry/model not set");
  if (ctx->model.kind == IPGPU_MODEL_IHARM && !ctx->d_dump)
    return set_err(ctx, IPGPU_ERR_STATE, "no fluid dump uploaded");
  const ipgpu_image &P = *img;
  if (P.nx <= 0 || P.ny <= 0 || P.maxnstep < 2) return set_err(ctx, IPGPU_ERR_ARG, "bad image size or maxnstep");
  if (P.emission_type != 4 && P.emission_type != 1)
    return set_err(ctx, IPGPU_ERR_ARG, "emission_type must be 4 (Dexter thermal) or 1 (Pandya thermal)");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  batch_evict(ctx);
  release_overflow(ctx);
  if (P.nx > 32767 || P.ny > 32767) return set_err(ctx, IPGPU_ERR_ARG, "image side above 32767 pixels");
  int tiles = ((P.nx + TILE - 1) / TILE) * ((P.ny + TILE - 1) / TILE);
  int nrays = tiles * WAVE;
  size_t npix = (size_t)P.nx * P.ny;
  int rc;
  if ((rc = grow(ctx, ctx->d_image, ctx->cap_image, npix))) return rc;
  if ((rc = grow(ctx, ctx->d_taus, ctx->cap_taus, npix))) return rc;
  if ((rc = grow(ctx, ctx->d_nstep_pix, ctx->cap_nstep_pix, npix))) return rc;
  if ((rc = grow(ctx, ctx->d_flag_pix, ctx->cap_flag_pix, npix))) return rc;
  if ((rc = grow(ctx, ctx->d_imageS, ctx->cap_imageS, npix * 5))) return rc;

  ctx->out_ny = P.ny;
  if ((rc = render_image(ctx, P, nrays))) return rc;
  ctx->last_npix = (int)npix;
  ctx->last_unpol = P.only_unpolarized;
  ctx->have_result = 1;
  return ipgpu_fetch(ctx, image, imageS, taus, nstep, flag);
}

/* The rays of ctx->pix (or the image's tiles) through the streamed or the
 * one-pass pipeline into the device image buffers (rows of ctx->out_ny).
 * nrays is a multiple of WAVE; the image buffers are sized by the caller. */
static int render_rays(ipgpu_ctx *ctx, const ipgpu_image &P, int nrays)
{
  int rc;
  /* mixed precision: the packet kernels gather the FP32 copy of the cell
   * records, written here from the dump as it is (the context forgets it
   * when this render returns) */
  struct D32Scope {
    ipgpu_ctx *c;
    ~D32Scope() { c->r_d32 = nullptr; }
  } d32_scope{ctx};
  ctx->r_d32 = nullptr;
  if (P.transfer_fp32 && ctx->model.kind == IPGPU_MODEL_IHARM && ctx->d_dump) {
    if ((rc = ensure_dump32(ctx))) return rc;
    const long long nc = (long long)ctx->dump_cells;
    k_dump_f32<<<(unsigned)((nc + 255) / 256), 256, 0, ctx->stream>>>(nc, ctx->d_dump, ctx->d_dump32);
    HIPCHK(ctx, hipGetLastError());
    ctx->r_d32 = ctx->d_dump32;
  }
  if ((rc = grow(ctx, ctx->d_nstep_ray, ctx->cap_nstep_ray, (size_t)nrays))) return rc;
  if ((rc = grow(ctx, ctx->d_flag_ray, ctx->cap_flag_ray, (size_t)nrays))) return rc;
  if ((rc = grow(ctx, ctx->d_out7, ctx->cap_out7, (size_t)nrays * 7))) return rc;
  int cap = ctx->traj_cap_steps < P.maxnstep ? ctx->traj_cap_steps : P.maxnstep;
  const int nreg = pool_regions(ctx, P, nrays);
  if (nreg > 0) cap = ctx->trace_budget + 1;   /* slots for the bounded phases' records */
  const bool two = P.target_nturns < 0 && ctx->trace_budget > 0 && (nreg > 0 || cap > ctx->trace_budget + 1);
  ctx->last_overflow = 0;
  if (two && (rc = render_streamed(ctx, P, nrays, cap, nreg))) return rc;
  const bool one = !two;
  if (one) {
    int stride;
    unsigned long long total;
    if ((rc = prepare_packets(ctx, P, nrays, false, cap, total, stride))) return rc;
    /* long rays of a polarized image: their chains run as composed segments */
    ipgpu_ctx::SegSet &S = ctx->segset[0];
    S.off.clear();
    if (stride == PK && !P.stokes_floors) {
      std::vector<int> &ns = ctx->h_ns, &list = ctx->h_list, &base = ctx->h_base;
      std::vector<long long> &off = ctx->h_off;
      ns.resize(nrays);
      HIPCHK(ctx, hipMemcpyAsync(ns.data(), ctx->d_nstep_ray, nrays * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      list.resize(nrays);
      off.assign((size_t)nrays + 1, 0);
      base.assign(nrays, -1);
      for (int r = 0; r < nrays; r++) {
        list[r] = r;
        off[r + 1] = off[r] + ns[r];   /* the packet offsets k_scan forms */
        off[r] *= PK;                  /* in doubles */
      }
      build_segments(S, list.data(), nrays, off.data(), ns, ctx->seg_min_steps, ctx->seg_len, P.maxnstep,
                     base.data());
      if (!S.off.empty()) {
        if ((rc = grow(ctx, ctx->d_segbase, ctx->cap_segbase, (size_t)nrays))) return rc;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_segbase, base.data(), nrays * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        if ((rc = upload_segments(ctx, S, ctx->stream))) return rc;
        long long lanes = (long long)S.off.size() * SEGCOL;
        k_seg<1><<<(unsigned)((lanes + WAVE - 1) / WAVE), WAVE, 0, ctx->stream>>>((int)S.off.size(), S.d_off,
                                                                                  S.d_q0, S.d_q1, ctx->d_pkt,
                                                                                  S.d_map, S.d_flag, 0);
        HIPCHK(ctx, hipGetLastError());
      }
    }
    const bool segd = !S.off.empty();
    if (segd) {
      k_chain<true><<<nrays / WAVE, WAVE, 0, ctx->stream>>>(ctx->geom, ctx->model, P, ctx->pix, nrays, cap, stride, ctx->d_off,
                                                             ctx->d_nstep_ray, ctx->d_traj, ctx->d_dump, ctx->d_pkt,
                                                             ctx->d_segbase, ctx->seg_len, S.d_map, S.d_flag,
                                                             ctx->d_out7, ctx->d_flag_ray, ctx->d_status, ctx->d_counts);
      HIPCHK(ctx, hipGetLastError());
    }
    k_chain<false><<<nrays / WAVE, WAVE, 0, ctx->stream>>>(ctx->geom, ctx->model, P, ctx->pix, nrays, cap, stride, ctx->d_off,
                                                            ctx->d_nstep_ray, ctx->d_traj, ctx->d_dump, ctx->d_pkt,
                                                            segd ? ctx->d_segbase : nullptr, ctx->seg_len, S.d_map,
                                                            S.d_flag, ctx->d_out7, ctx->d_flag_ray, ctx->d_status,
                                                            ctx->d_counts);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  }
  const int e1 = one ? 4 : 22;
  double f3 = pow(P.freqcgs, 3);
  k_finalize<<<(nrays + 255) / 256, 256, 0, ctx->stream>>>(P, ctx->pix, ctx->out_ny, nrays, f3, ctx->d_out7, ctx->d_flag_ray,
                                                           ctx->d_nstep_ray, ctx->d_image, ctx->d_imageS,
                                                           ctx->d_taus, ctx->d_nstep_pix, ctx->d_flag_pix,
                                                           ctx->d_status);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev[e1], ctx->stream));
  int status[NSTATUS];
  HIPCHK(ctx, hipMemcpyAsync(status, ctx->d_status, sizeof(status), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->counts, ctx->d_counts, sizeof(ctx->counts), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (one) {
    /* trace, packets, chains, finalize */
    for (int q = 0; q < 4; q++) hipEventElapsedTime(&ctx->ms_stage[q], ctx->ev[q], ctx->ev[q + 1]);
    ctx->ms_geo = ctx->ms_stage[0];
    ctx->ms_xfer = ctx->ms_stage[1] + ctx->ms_stage[2];
  } else {
    /* [0] the trace launches, [1] the span of the finished rays' transfer
     * (their own streams, concurrent with later trace phases), [2] what
     * follows the last trace launch until the streams have joined (the
     * longest rays' transfer, the wait), [3] finalize */
    const int K = ctx->nphase;
    float a, tr = 0.f;
    for (int k = 0; k <= K; k++) {
      hipEventElapsedTime(&a, ctx->ev[k == 0 ? 0 : 2 * k], ctx->ev[2 * k + 1]);
      tr += a;
    }
    ctx->ms_stage[0] = tr;
    float f1 = 0.f;
    for (int k = 0; k < K; k++) {
      hipEventElapsedTime(&a, ctx->ev[19], ctx->ev[15 + k]);
      f1 = a > f1 ? a : f1;
    }
    ctx->ms_stage[1] = f1;
    hipEventElapsedTime(&ctx->ms_stage[2], ctx->ev[2 * K + 1], ctx->ev[21]);
    hipEventElapsedTime(&ctx->ms_stage[3], ctx->ev[21], ctx->ev[22]);
    hipEventElapsedTime(&ctx->ms_geo, ctx->ev[0], ctx->ev[2 * K + 1]);
    ctx->ms_xfer = ctx->ms_stage[2];
  }
  ctx->streamed = !one;
  hipEventElapsedTime(&ctx->ms_total, ctx->ev[0], ctx->ev[e1]);
  if (!one) {
    if (status[0] != 0)
      return set_err(ctx, IPGPU_ERR_STATE, "continued trace of a ray past the record capacity differs");
    /* the next render's record capacity: this one's longest ray, if the
     * records of a chunk of rays then stay within half the render memory */
    const long long want = (long long)ctx->last_max_steps + 1;
    const long long lim = (long long)(ctx->render_bytes / 2 / ((size_t)nrays * REC * sizeof(double)));
    if (want > ctx->traj_cap_steps && want <= lim) ctx->traj_cap_steps = (int)want;
  }
  if (status[1] != 0x7fffffff) {
    int p = status[1] - 1;
    ctx->bad_i = p >> 16;
    ctx->bad_j = p & 0xffff;
    snprintf(ctx->err, sizeof(ctx->err), "MaxNStep exceeded in pixel %d %d", ctx->bad_i, ctx->bad_j);
    return IPGPU_ERR_MAXNSTEP;
  }
  if (status[2] != 0x7fffffff) {
    int p = status[2] - 1;
    ctx->bad_i = p >> 16;
    ctx->bad_j = p & 0xffff;
    snprintf(ctx->err, sizeof(ctx->err), "NaN in image at pixel %d %d", ctx->bad_i, ctx->bad_j);
    return IPGPU_ERR_NAN;
  }
  return IPGPU_OK;
}

/* ------------------------------------------------------ adaptive refinement */

static int adaptive_fetch(ipgpu_ctx *ctx, const ipgpu_image &P, int nyi, double *image, double *imageS,
                          double *taus, int *nstep, int *flag, int *interp);

int ipgpu_render_adaptive(ipgpu_ctx *ctx, const ipgpu_image *img, const ipgpu_refine *rf, double *image,
                          double *imageS, double *taus, int *nstep, int *flag, int *interp)
{
  if (!ctx || !img || !rf || !image) return IPGPU_ERR_ARG;
  batch_evict(ctx);
  release_overflow(ctx);
  const ipgpu_image &P = *img;
  if (rf->nx_min <= 0 || rf->nx_min == P.nx) {
    ctx->ref_levels = 1;
    ctx->ref_iavg = 0.;
    int rc = ipgpu_render(ctx, img, image, imageS, taus, nstep, flag);
    ctx->ref_traced[0] = (long long)P.nx * P.ny;
    ctx->ref_interp[0] = 0;
    if (rc == IPGPU_OK && interp) memset(interp, 0, sizeof(int) * (size_t)P.nx * P.ny);
    return rc;
  }
  if (!ctx->has_geom || !ctx->has_model) return set_err(ctx, IPGPU_ERR_STATE, "geometry/model not set");
  if (ctx->model.kind == IPGPU_MODEL_IHARM && !ctx->d_dump)
    return set_err(ctx, IPGPU_ERR_STATE, "no fluid dump uploaded");
  if (P.nx <= 0 || P.ny <= 0 || P.maxnstep < 2 || rf->ny_min <= 0 || rf->nx_min > P.nx || rf->ny_min > P.ny)
    return set_err(ctx, IPGPU_ERR_ARG, "bad image size, refinement size or maxnstep");
  if (P.emission_type != 4 && P.emission_type != 1)
    return set_err(ctx, IPGPU_ERR_ARG, "emission_type must be 4 (Dexter thermal) or 1 (Pandya thermal)");
  /* src/main.c:104-124 */
  const int ratio = (P.nx - P.nx % 2) / (rf->nx_min - rf->nx_min % 2);
  if (ratio < 1) return set_err(ctx, IPGPU_ERR_ARG, "nx_min too large");
  const int levels = (int)(log2((double)ratio) + 1);
  if (levels > IPGPU_MAX_LEVELS) return set_err(ctx, IPGPU_ERR_ARG, "too many refinement levels");
  const int nxi = (levels > 1 && P.nx % 2 == 0) ? P.nx + 1 : P.nx;
  const int nyi = (levels > 1 && P.ny % 2 == 0) ? P.ny + 1 : P.ny;
  const int nxmi = (levels > 1 && P.nx % 2 == 0) ? rf->nx_min + 1 : rf->nx_min;
  const int nymi = (levels > 1 && P.ny % 2 == 0) ? rf->ny_min + 1 : rf->ny_min;
  if (levels == 1) {
    ipgpu_refine r1 = *rf;
    r1.nx_min = -1;
    return ipgpu_render_adaptive(ctx, img, &r1, image, imageS, taus, nstep, flag, interp);
  }
  const int sp = 1 << (levels - 1);
  if (nxmi < 2 || nymi < 2 || nxi - 1 != (nxmi - 1) * sp || nyi - 1 != (nymi - 1) * sp || nxi > 32767
      || nyi > 32767)
    return set_err(ctx, IPGPU_ERR_ARG, "the internal grid must be (n_min - 1) 2^(levels - 1) + 1 on both axes");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const size_t npi = (size_t)nxi * nyi;
  int rc;
  if ((rc = grow(ctx, ctx->d_image, ctx->cap_image, npi))) return rc;
  if ((rc = grow(ctx, ctx->d_taus, ctx->cap_taus, npi))) return rc;
  if ((rc = grow(ctx, ctx->d_nstep_pix, ctx->cap_nstep_pix, npi))) return rc;
  if ((rc = grow(ctx, ctx->d_flag_pix, ctx->cap_flag_pix, npi))) return rc;
  if ((rc = grow(ctx, ctx->d_imageS, ctx->cap_imageS, npi * 5))) return rc;
  if ((rc = grow(ctx, ctx->d_interp, ctx->cap_interp, npi))) return rc;
  if ((rc = grow(ctx, ctx->d_rlist, ctx->cap_rlist, npi + WAVE))) return rc;
  if (!ctx->d_rcount) {
    HIPCHK(ctx, hipMalloc((void **)&ctx->d_rcount, 2 * sizeof(unsigned long long)));
    HIPCHK(ctx, hipMalloc((void **)&ctx->d_iavg, sizeof(double)));
  }
  hipStream_t st = ctx->stream;
  HIPCHK(ctx, hipMemsetAsync(ctx->d_image, 0, npi * sizeof(double), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_taus, 0, npi * sizeof(double), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_imageS, 0, npi * 5 * sizeof(double), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_nstep_pix, 0, npi * sizeof(int), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_flag_pix, 0, npi * sizeof(int), st));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_interp, 0, npi * sizeof(int), st));
  ctx->have_result = 0;       /* the device buffers hold the internal grid, not ipgpu_fetch's layout */
  ctx->ref_levels = levels;
  for (int l = 0; l < IPGPU_MAX_LEVELS; l++) ctx->ref_traced[l] = ctx->ref_interp[l] = 0;
  const int nbase = nxmi * nymi;
  RefineK R{nxi, nyi, sp, nymi, rf->refine_abs, rf->refine_rel, rf->refine_cut, rf->nearest_neighbor};
  const double f3 = pow(P.freqcgs, 3);
  unsigned int *cnt = reinterpret_cast<unsigned int *>(ctx->d_rcount);
  /* Speculative: every pixel of the internal grid in one render, then each
   * level's decisions and interpolations in place.  A pixel's result does
   * not depend on which pixels are traced, so the image is the one tracing
   * level by level gives, for one render's latency instead of one per
   * level (the levels' lists are short and their rays long).  Above
   * refine_spec_max pixels, or when a pixel the levels would not trace
   * fails (maxnstep, NaN), the levels' lists are traced instead. */
  if ((long long)npi <= ctx->refine_spec_max) {
    const int npad = (int)((npi + WAVE - 1) / WAVE * WAVE);
    {
      std::vector<int> hl(npad, -1);
      for (int i = 0; i < nxi; i++)
        for (int j = 0; j < nyi; j++) hl[(size_t)i * nyi + j] = (i << 16) | j;
      HIPCHK(ctx, hipMemcpyAsync(ctx->d_rlist, hl.data(), npad * sizeof(int), hipMemcpyHostToDevice, st));
      HIPCHK(ctx, hipStreamSynchronize(st));
    }
    ctx->pix = ctx->d_rlist;
    ctx->out_ny = nyi;
    rc = render_rays(ctx, P, npad);
    ctx->pix = nullptr;
    if (rc == IPGPU_OK) {
      ctx->ref_traced[0] = nbase;
      k_refine_iavg<<<1, WAVE, 0, st>>>(R, nbase, 1, f3, ctx->d_out7, ctx->d_iavg);
      HIPCHK(ctx, hipGetLastError());
      HIPCHK(ctx, hipMemcpyAsync(&ctx->ref_iavg, ctx->d_iavg, sizeof(double), hipMemcpyDeviceToHost, st));
      long long interp_total = 0;
      for (int lev = 1; lev < levels; lev++) {
        const int ns = sp >> lev;
        const long long cand = (long long)((nxi - 1) / ns + 1) * ((nyi - 1) / ns + 1);
        HIPCHK(ctx, hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned int), st));
        k_refine_level<<<(unsigned)((cand + 255) / 256), 256, 0, st>>>(
            R, ns, P.only_unpolarized, ctx->d_iavg, ctx->d_image, ctx->d_imageS, ctx->d_taus, ctx->d_nstep_pix,
            ctx->d_flag_pix, ctx->d_interp, nullptr, cnt);
        HIPCHK(ctx, hipGetLastError());
        unsigned int hc[2];
        HIPCHK(ctx, hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, st));
        HIPCHK(ctx, hipStreamSynchronize(st));
        interp_total += hc[1];
        ctx->ref_traced[lev] = hc[0];
        ctx->ref_interp[lev] = interp_total;
      }
      return adaptive_fetch(ctx, P, nyi, image, imageS, taus, nstep, flag, interp);
    }
    if (rc != IPGPU_ERR_MAXNSTEP && rc != IPGPU_ERR_NAN) return rc;
    HIPCHK(ctx, hipMemsetAsync(ctx->d_image, 0, npi * sizeof(double), st));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_taus, 0, npi * sizeof(double), st));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_imageS, 0, npi * 5 * sizeof(double), st));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_nstep_pix, 0, npi * sizeof(int), st));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_flag_pix, 0, npi * sizeof(int), st));
  }
  /* coarse pass: the base grid in the reference's loop order (i outer) */
  const int nbpad = (nbase + WAVE - 1) / WAVE * WAVE;
  {
    std::vector<int> hl(nbpad, -1);
    for (int a = 0; a < nxmi; a++)
      for (int b = 0; b < nymi; b++) hl[(size_t)a * nymi + b] = ((a * sp) << 16) | (b * sp);
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_rlist, hl.data(), nbpad * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
  }
  ctx->pix = ctx->d_rlist;
  ctx->out_ny = nyi;
  rc = render_rays(ctx, P, nbpad);
  ctx->pix = nullptr;
  if (rc) return rc;
  ctx->ref_traced[0] = nbase;
  k_refine_iavg<<<1, WAVE, 0, st>>>(R, nbase, 0, f3, ctx->d_out7, ctx->d_iavg);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(&ctx->ref_iavg, ctx->d_iavg, sizeof(double), hipMemcpyDeviceToHost, st));
  long long interp_total = 0;
  for (int lev = 1; lev < levels; lev++) {
    const int ns = sp >> lev;
    const long long cand = (long long)((nxi - 1) / ns + 1) * ((nyi - 1) / ns + 1);
    HIPCHK(ctx, hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned int), st));
    k_refine_level<<<(unsigned)((cand + 255) / 256), 256, 0, st>>>(
        R, ns, P.only_unpolarized, ctx->d_iavg, ctx->d_image, ctx->d_imageS, ctx->d_taus, ctx->d_nstep_pix,
        ctx->d_flag_pix, ctx->d_interp, ctx->d_rlist, cnt);
    HIPCHK(ctx, hipGetLastError());
    unsigned int hc[2];
    HIPCHK(ctx, hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    const int n = (int)hc[0];
    interp_total += hc[1];
    ctx->ref_traced[lev] = n;
    ctx->ref_interp[lev] = interp_total;
    if (n == 0) continue;
    const int npad = (n + WAVE - 1) / WAVE * WAVE;
    if (npad > n) HIPCHK(ctx, hipMemsetAsync(ctx->d_rlist + n, 0xff, (size_t)(npad - n) * sizeof(int), st));
    ctx->pix = ctx->d_rlist;
    rc = render_rays(ctx, P, npad);
    ctx->pix = nullptr;
    if (rc) return rc;
  }
  return adaptive_fetch(ctx, P, nyi, image, imageS, taus, nstep, flag, interp);
}

/* the nx x ny part of the internal grid (rows of nyi) */
static int adaptive_fetch(ipgpu_ctx *ctx, const ipgpu_image &P, int nyi, double *image, double *imageS,
                          double *taus, int *nstep, int *flag, int *interp)
{
  /* each array whole into host memory in one copy, then its rows: strided
   * copies into pageable memory go row by row */
  hipStream_t st = ctx->stream;
  const size_t nxi = (size_t)P.nx + ((ctx->ref_levels > 1 && P.nx % 2 == 0) ? 1 : 0);
  const size_t npi = nxi * nyi;
  std::vector<double> hd(npi * 5);
  std::vector<int> hi(npi);
  auto rows = [&](const void *src, void *dst, size_t elem, size_t per) {
    for (size_t i = 0; i < (size_t)P.nx; i++)
      memcpy((char *)dst + i * P.ny * per * elem, (const char *)src + i * nyi * per * elem, P.ny * per * elem);
  };
  auto getd = [&](const double *dsrc, double *dst, size_t per) -> int {
    HIPCHK(ctx, hipMemcpyAsync(hd.data(), dsrc, npi * per * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    rows(hd.data(), dst, sizeof(double), per);
    return 0;
  };
  auto geti = [&](const int *dsrc, int *dst) -> int {
    HIPCHK(ctx, hipMemcpyAsync(hi.data(), dsrc, npi * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    rows(hi.data(), dst, sizeof(int), 1);
    return 0;
  };
  int rc;
  if ((rc = getd(ctx->d_image, image, 1))) return rc;
  if (taus && (rc = getd(ctx->d_taus, taus, 1))) return rc;
  if (imageS && !P.only_unpolarized && (rc = getd(ctx->d_imageS, imageS, 5))) return rc;
  if (nstep && (rc = geti(ctx->d_nstep_pix, nstep))) return rc;
  if (flag && (rc = geti(ctx->d_flag_pix, flag))) return rc;
  if (interp && (rc = geti(ctx->d_interp, interp))) return rc;
  return IPGPU_OK;
}

int ipgpu_adaptive_stats(const ipgpu_ctx *ctx, int *levels, double *Iavg, long long *traced,
                         long long *interpolated)
{
  if (!ctx) return IPGPU_ERR_ARG;
  if (levels) *levels = ctx->ref_levels;
  if (Iavg) *Iavg = ctx->ref_iavg;
  for (int l = 0; l < IPGPU_MAX_LEVELS; l++) {
    if (traced) traced[l] = ctx->ref_traced[l];
    if (interpolated) interpolated[l] = ctx->ref_interp[l];
  }
  return IPGPU_OK;
}

int ipgpu_render_noalloc(ipgpu_ctx *ctx, const ipgpu_image *img, double *image, double *imageS, double *taus,
                         int *nstep, int *flag)
{
  if (!ctx || !img) return IPGPU_ERR_ARG;
  if (!ctx->has_geom || !ctx->has_model) return set_err(ctx, IPGPU_ERR_STATE, "geometry/model not set");
  if (ctx->model.kind == IPGPU_MODEL_IHARM && !ctx->d_dump)
    return set_err(ctx, IPGPU_ERR_STATE, "no fluid dump uploaded");
  const ipgpu_image &P = *img;
  if (P.nx <= 0 || P.ny <= 0 || P.maxnstep < 2) return set_err(ctx, IPGPU_ERR_ARG, "bad image size or maxnstep");
  if (P.nx > 32767 || P.ny > 32767) return set_err(ctx, IPGPU_ERR_ARG, "image side above 32767 pixels");
  if (P.emission_type != 4 && P.emission_type != 1)
    return set_err(ctx, IPGPU_ERR_ARG, "emission_type must be 4 (Dexter thermal) or 1 (Pandya thermal)");
  if (P.target_nturns >= 0)
    return set_err(ctx, IPGPU_ERR_ARG, "ipgpu_render_noalloc: subring selection (target_nturns >= 0) not supported");
  if (!ctx->d_scratch) return set_err(ctx, IPGPU_ERR_STATE, "scratch address not set up");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int tiles = ((P.nx + TILE - 1) / TILE) * ((P.ny + TILE - 1) / TILE);
  const long long nrays = (long long)tiles * WAVE;
  /* layout: status (byte 0), counters (64), step counts (256), offsets,
   * then the records */
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t o_ns = 256, o_off = up(o_ns + (size_t)nrays * sizeof(int));
  const size_t o_arena = up(o_off + ((size_t)nrays + 1) * sizeof(long long));
  const long long rec_bytes = REC * (long long)sizeof(double);
  const long long A = o_arena < IPGPU_SCRATCH_BYTES ? (long long)((IPGPU_SCRATCH_BYTES - o_arena) / rec_bytes) : 0;
  const long long Ap = A - ((long long)P.maxnstep + 1);
  if (Ap < (long long)P.maxnstep + 1) {
    snprintf(ctx->err, sizeof(ctx->err),
             "ipgpu_render_noalloc: %dx%d image with maxnstep %d needs more than the %zu-byte scratch "
             "(per-ray tables of %zu bytes and room for two rays of maxnstep records)",
             P.nx, P.ny, P.maxnstep, (size_t)IPGPU_SCRATCH_BYTES, o_arena);
    return IPGPU_ERR_ARG;
  }
  char *base = (char *)ctx->d_scratch;
  int *d_status = (int *)base;
  unsigned long long *d_counts = (unsigned long long *)(base + 64);
  int *d_ns = (int *)(base + o_ns);
  long long *d_off = (long long *)(base + o_off);
  double *arena = (double *)(base + o_arena);
  hipStream_t st = ctx->stream;
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], st));
  k_na_init<<<1, 64, 0, st>>>(d_status, d_counts);
  HIPCHK(ctx, hipGetLastError());
  k_na_count<<<(unsigned)(2 * nrays / WAVE), WAVE, 0, st>>>(ctx->geom, P, (int)nrays, d_ns, d_counts);
  HIPCHK(ctx, hipGetLastError());
  k_scan<<<1, SCAN_THREADS, 0, st>>>((int)nrays, d_ns, d_off);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], st));
  long long total = 0;
  HIPCHK(ctx, hipMemcpyAsync(&total, d_off + nrays, sizeof(long long), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  const long long nchunks = (total + nrays + Ap - 1) / Ap;
  auto kx = P.transfer_fp32 ? k_na_xfer<true> : k_na_xfer<false>;
  for (long long c = 0; c < nchunks; c++) {
    k_na_store<<<(unsigned)(2 * nrays / WAVE), WAVE, 0, st>>>(ctx->geom, P, (int)nrays, c, Ap, d_ns, d_off, arena,
                                                              d_status);
    HIPCHK(ctx, hipGetLastError());
    kx<<<(unsigned)(nrays / WAVE), WAVE, 0, st>>>(ctx->geom, ctx->model, P, (int)nrays, c, Ap, d_ns, d_off, arena,
                                                   ctx->d_dump, image, P.only_unpolarized ? nullptr : imageS, taus,
                                                   nstep, flag, d_status, d_counts);
    HIPCHK(ctx, hipGetLastError());
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], st));
  int status[NSTATUS];
  HIPCHK(ctx, hipMemcpyAsync(status, d_status, sizeof(status), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipMemcpyAsync(ctx->counts, d_counts, sizeof(ctx->counts), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  hipEventElapsedTime(&ctx->ms_geo, ctx->ev[0], ctx->ev[1]);
  hipEventElapsedTime(&ctx->ms_total, ctx->ev[0], ctx->ev[2]);
  ctx->ms_xfer = ctx->ms_total - ctx->ms_geo;
  ctx->last_chunks = (int)nchunks;
  ctx->last_overflow = 0;
  if (status[0] != 0) return set_err(ctx, IPGPU_ERR_STATE, "ipgpu_render_noalloc: store pass step count differs");
  if (status[1] != 0x7fffffff) {
    const int p = status[1] - 1;
    ctx->bad_i = p >> 16;
    ctx->bad_j = p & 0xffff;
    snprintf(ctx->err, sizeof(ctx->err), "MaxNStep exceeded in pixel %d %d", ctx->bad_i, ctx->bad_j);
    return IPGPU_ERR_MAXNSTEP;
  }
  if (status[2] != 0x7fffffff) {
    const int p = status[2] - 1;
    ctx->bad_i = p >> 16;
    ctx->bad_j = p & 0xffff;
    snprintf(ctx->err, sizeof(ctx->err), "NaN in image at pixel %d %d", ctx->bad_i, ctx->bad_j);
    return IPGPU_ERR_NAN;
  }
  return IPGPU_OK;
}

int ipgpu_fetch(ipgpu_ctx *ctx, double *image, double *imageS, double *taus, int *nstep, int *flag)
{
  if (!ctx) return IPGPU_ERR_ARG;
  if (!ctx->have_result) return set_err(ctx, IPGPU_ERR_STATE, "no render result");
  size_t npix = ctx->last_npix;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (image) HIPCHK(ctx, hipMemcpyAsync(image, ctx->d_image, npix * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (imageS && !ctx->last_unpol)
    HIPCHK(ctx, hipMemcpyAsync(imageS, ctx->d_imageS, npix * 5 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (taus) HIPCHK(ctx, hipMemcpyAsync(taus, ctx->d_taus, npix * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (nstep) HIPCHK(ctx, hipMemcpyAsync(nstep, ctx->d_nstep_pix, npix * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (flag) HIPCHK(ctx, hipMemcpyAsync(flag, ctx->d_flag_pix, npix * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IPGPU_OK;
}

/* ------------------------------------------------------------ batched render */

static bool same_trace(const ipgpu_image &a, const ipgpu_image &b)
{
  for (int k = 0; k < 4; k++)
    if (a.Xcam[k] != b.Xcam[k]) return false;
  return a.nx == b.nx && a.ny == b.ny && a.fovx == b.fovx && a.fovy == b.fovy && a.rotcam == b.rotcam
      && a.xoff == b.xoff && a.yoff == b.yoff && a.eps == b.eps && a.maxnstep == b.maxnstep
      && a.old_centering == b.old_centering && a.target_nturns == b.target_nturns
      && a.subring_dtheta == b.subring_dtheta && a.freq > 0. && b.freq > 0.;
}

/* Every field of the image the same (the struct's padding aside). */
static bool same_image(const ipgpu_image &a, const ipgpu_image &b)
{
  return same_trace(a, b) && a.freq == b.freq && a.freqcgs == b.freqcgs
      && a.only_unpolarized == b.only_unpolarized && a.qu_conv == b.qu_conv && a.stokes_floors == b.stokes_floors
      && a.counterjet == b.counterjet && a.emission_type == b.emission_type && a.transfer_fp32 == b.transfer_fp32;
}

static int group_of_host(const std::vector<BGroup> &groups, long long r)
{
  int lo = 0, hi = (int)groups.size();
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (groups[mid].ray_base <= r) lo = mid;
    else hi = mid;
  }
  return lo;
}

/* The transfer of a batch's nitems work items (d_items, d_jobs, records at
 * d_traj / traj2), on the context's stream: the persistent work-queue
 * launch or one wave per 64 items.  clear_status: first reset the status
 * words and the radiating step count on the device. */
static int batch_transfer(ipgpu_ctx *ctx, bool f32, bool multi, int nitems, const double *traj2,
                          const int *nturns, bool clear_status)
{
  hipStream_t st = ctx->stream;
  if (clear_status || (ctx->batch_queue && nitems > 0)) {
    k_queue_reset<<<1, WAVE, 0, st>>>(ctx->d_queue, clear_status ? ctx->d_bstatus : nullptr,
                                      clear_status ? ctx->d_counts : nullptr, clear_status ? ctx->d_bcounts : nullptr);
    HIPCHK(ctx, hipGetLastError());
  }
  if (nitems <= 0) return 0;
  /* mixed precision: the FP32 copy of the records, from the dump as it is */
  const float *d32 = nullptr;
  if (f32 && ctx->d_dump32 && ctx->dump32_cells == ctx->dump_cells && ctx->model.kind == IPGPU_MODEL_IHARM) {
    const long long nc = (long long)ctx->dump_cells;
    k_dump_f32<<<(unsigned)((nc + 255) / 256), 256, 0, st>>>(nc, ctx->d_dump, ctx->d_dump32);
    HIPCHK(ctx, hipGetLastError());
    d32 = ctx->d_dump32;
  }
  if (multi) {
    /* items (ray, job subset) */
    auto km = f32 ? k_xfer_multi<true> : k_xfer_multi<false>;
    const int nw = ctx->batch_queue_waves > 0 ? ctx->batch_queue_waves : ctx->xm_occ[f32 ? 1 : 0] * ctx->ncu;
    const unsigned grid = ctx->batch_queue ? (unsigned)nw : (unsigned)((nitems + WAVE - 1) / WAVE);
    km<<<grid, WAVE, 0, st>>>(ctx->geom, ctx->model, ctx->d_jobs, ctx->d_groups, ctx->d_items, ctx->d_subsets,
                              nitems, ctx->batch_queue ? ctx->d_queue : nullptr, ctx->d_off, ctx->d_traj, traj2,
                              ctx->d_nstep_ray, nturns, ctx->d_dump, ctx->d_bimage, ctx->d_bimageS, ctx->d_btaus,
                              ctx->d_bnstep, ctx->d_bflag, ctx->d_bstatus, ctx->d_counts, ctx->d_bcounts, d32);
    HIPCHK(ctx, hipGetLastError());
    return 0;
  }
  if (ctx->batch_queue) {
    const int nw = ctx->batch_queue_waves > 0 ? ctx->batch_queue_waves : ctx->xq_occ[f32 ? 1 : 0] * ctx->ncu;
    auto kq = f32 ? k_xfer_queue<true> : k_xfer_queue<false>;
    kq<<<nw, WAVE, 0, st>>>(ctx->geom, ctx->model, ctx->d_jobs, ctx->d_groups, ctx->d_items, nitems, ctx->d_queue,
                            ctx->d_off, ctx->d_traj, traj2, ctx->d_nstep_ray, nturns, ctx->d_dump, ctx->d_bimage,
                            ctx->d_bimageS, ctx->d_btaus, ctx->d_bnstep, ctx->d_bflag, ctx->d_bstatus, ctx->d_counts,
                            d32);
  } else {
    auto kx = f32 ? k_xfer_batch<true> : k_xfer_batch<false>;
    kx<<<(nitems + WAVE - 1) / WAVE, WAVE, 0, st>>>(ctx->geom, ctx->model, ctx->d_jobs, ctx->d_groups, ctx->d_items,
                                                    nitems, ctx->d_off, ctx->d_traj, traj2, ctx->d_nstep_ray, nturns,
                                                    ctx->d_dump, ctx->d_bimage, ctx->d_bimageS, ctx->d_btaus,
                                                    ctx->d_bnstep, ctx->d_bflag, ctx->d_bstatus, ctx->d_counts,
                                                    d32);
  }
  HIPCHK(ctx, hipGetLastError());
  return 0;
}

static void drop_batch_graph(ipgpu_ctx *ctx)
{
  if (ctx->bgraph_exec) hipGraphExecDestroy(ctx->bgraph_exec);
  if (ctx->bgraph) hipGraphDestroy(ctx->bgraph);
  ctx->bgraph_exec = nullptr;
  ctx->bgraph = nullptr;
}

/* Every entry point that reuses the ray buffers (records, offsets, step
 * counts) or changes the geometry ends the resident batch. */
static void batch_evict(ipgpu_ctx *ctx)
{
  ctx->blast.valid = false;
  drop_batch_graph(ctx);
}

/* The batch's first-pass overflow arena and its tables are scratch of the
 * batch's traces: a single-image render or an adjoint gives their memory
 * back (the next batch allocates them again). */
static void release_overflow(ipgpu_ctx *ctx)
{
  if (!ctx->d_ovf_arena && !ctx->d_ovf_table && !ctx->d_ovf_ok) return;
  if (hipSetDevice(ctx->device) != hipSuccess) return;
  if (ctx->d_ovf_arena && hipFree(ctx->d_ovf_arena) == hipSuccess)
    ctx->dev_bytes -= ctx->ovf_arena_doubles * sizeof(double);
  if (ctx->d_ovf_table && hipFree(ctx->d_ovf_table) == hipSuccess) ctx->dev_bytes -= ctx->cap_ovf_table * sizeof(int);
  if (ctx->d_ovf_ok && hipFree(ctx->d_ovf_ok) == hipSuccess) ctx->dev_bytes -= ctx->cap_ovf_ok * sizeof(int);
  ctx->d_ovf_arena = nullptr;
  ctx->ovf_arena_doubles = 0;
  ctx->d_ovf_table = nullptr;
  ctx->cap_ovf_table = 0;
  ctx->d_ovf_ok = nullptr;
  ctx->cap_ovf_ok = 0;
}

int ipgpu_render_batch(ipgpu_ctx *ctx, int njobs, const ipgpu_job *jobs)
{
  if (!ctx || njobs <= 0 || !jobs) return IPGPU_ERR_ARG;
  if (!ctx->has_geom || !ctx->has_model) return set_err(ctx, IPGPU_ERR_STATE, "geometry/model not set");
  if (ctx->model.kind == IPGPU_MODEL_IHARM && !ctx->d_dump)
    return set_err(ctx, IPGPU_ERR_STATE, "no fluid dump uploaded");
  for (int j = 0; j < njobs; j++) {
    const ipgpu_image &P = jobs[j].image;
    if (P.nx <= 0 || P.ny <= 0 || P.maxnstep < 2 || !(P.freq > 0.) || !(jobs[j].m_unit_scale > 0.))
      return set_err(ctx, IPGPU_ERR_ARG, "bad job: image size, maxnstep, frequency or m_unit_scale");
    if (P.emission_type != 4 && P.emission_type != 1)
      return set_err(ctx, IPGPU_ERR_ARG, "emission_type must be 4 or 1");
    if ((P.transfer_fp32 != 0) != (jobs[0].image.transfer_fp32 != 0))
      return set_err(ctx, IPGPU_ERR_ARG, "all jobs of a batch must use the same transfer precision");
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc;
  batch_evict(ctx);
  if (jobs[0].image.transfer_fp32 && (rc = ensure_dump32(ctx))) return rc;
  /* trace groups */
  std::vector<BGroup> groups;
  std::vector<BJob> bj(njobs);
  std::vector<std::vector<int>> members;
  long long pix_total = 0;
  bool turns = false;
  for (int j = 0; j < njobs; j++) {
    const ipgpu_image &P = jobs[j].image;
    int g = -1;
    for (size_t q = 0; q < groups.size(); q++)
      if (same_trace(groups[q].P, P)) {
        g = (int)q;
        break;
      }
    if (g < 0) {
      BGroup gr{};
      gr.P = P;
      gr.nrays = ((P.nx + TILE - 1) / TILE) * ((P.ny + TILE - 1) / TILE) * WAVE;
      gr.ray_base = groups.empty() ? 0 : groups.back().ray_base + groups.back().nrays;
      groups.push_back(gr);
      members.emplace_back();
      g = (int)groups.size() - 1;
    }
    members[g].push_back(j);
    BJob &b = bj[j];
    b.P = P;
    b.group = g;
    b.pix_off = pix_total;
    pix_total += (long long)P.nx * P.ny;
    double kf = P.freq / groups[g].P.freq;
    b.sc.kf = kf;
    b.sc.ikf = 1. / kf;
    b.sc.ne = jobs[j].m_unit_scale;
    b.sc.b = sqrt(jobs[j].m_unit_scale);
    if (P.target_nturns >= 0) turns = true;
  }
  const long long rays_total = groups.back().ray_base + groups.back().nrays;
  const int ngroups = (int)groups.size();
  /* the multi-job transfer: each group's jobs in subsets of at most MJ, one
   * work item per (ray, subset) */
  const bool multi = ctx->batch_multi != 0;
  std::vector<int4> subsets;
  std::vector<std::vector<int>> gsub(ngroups);
  if (multi) {
    for (int g = 0; g < ngroups; g++)
      for (size_t a = 0; a < members[g].size(); a += MJ) {
        const int n = (int)std::min((size_t)MJ, members[g].size() - a);
        int4 v;
        v.x = n;
        v.y = members[g][a];
        v.z = members[g][a + (n > 1 ? 1 : 0)];
        v.w = members[g][a + (n > 2 ? 2 : 0)];
        gsub[g].push_back((int)subsets.size());
        subsets.push_back(v);
      }
    if ((rc = grow(ctx, ctx->d_subsets, ctx->cap_subsets, subsets.size()))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_subsets, subsets.data(), subsets.size() * sizeof(int4), hipMemcpyHostToDevice,
                               ctx->stream));
  }
  /* a ray's items: (job, ray) each, or (ray, subset) in multi mode.  Items
   * go longest ray first (the launch's tail is then the short rays); in
   * multi mode rays in the same band of 2^ITEM_BAND steps have their items
   * ordered by subset size, so that a wave's lanes mostly run as many jobs.
   * A wave's lanes run in lockstep, each step with the largest subset among
   * its lanes still running: in a model of the 64-job set (items' step
   * counts and subset sizes, waves of 64 in this order) that costs 4.7%
   * over the items' own work with bands of 256 steps and 2.5% with 32, but
   * measured, 32-step bands made the FP64 transfer slower (357 -> 366 ms):
   * wider bands keep a wave's items on neighbouring rays of one trace
   * group, whose records and dump cells the caches then share; 512-step
   * bands measured the same as 256 */
  auto ray_items = [&](int r, int g, int sz, auto &&emit) {
    if (!multi) {
      if (sz == 1)
        for (int jb : members[g]) emit(make_int2(jb, r));
      return;
    }
    for (int sid : gsub[g])
      if (subsets[sid].x == sz) emit(make_int2(r, sid));
  };
  const int max_sz = multi ? MJ : 1;
  if ((rc = grow(ctx, ctx->d_groups, ctx->cap_groups, groups.size()))) return rc;
  if ((rc = grow(ctx, ctx->d_jobs, ctx->cap_jobs, (size_t)njobs))) return rc;
  if ((rc = grow(ctx, ctx->d_bimage, ctx->cap_bimage, (size_t)pix_total))) return rc;
  if ((rc = grow(ctx, ctx->d_bimageS, ctx->cap_bimageS, (size_t)pix_total * 5))) return rc;
  if ((rc = grow(ctx, ctx->d_btaus, ctx->cap_btaus, (size_t)pix_total))) return rc;
  if ((rc = grow(ctx, ctx->d_bnstep, ctx->cap_bnstep, (size_t)pix_total))) return rc;
  if ((rc = grow(ctx, ctx->d_bflag, ctx->cap_bflag, (size_t)pix_total))) return rc;
  if (!ctx->d_bstatus) HIPCHK(ctx, hipMalloc((void **)&ctx->d_bstatus, 2 * sizeof(unsigned long long)));
  if (!ctx->d_queue) HIPCHK(ctx, hipMalloc((void **)&ctx->d_queue, 64 * sizeof(unsigned int)));
  if (!ctx->d_bcounts) HIPCHK(ctx, hipMalloc((void **)&ctx->d_bcounts, 2 * sizeof(unsigned long long)));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_bcounts, 0, 2 * sizeof(unsigned long long), ctx->stream));
  for (long long &v : ctx->bstat) v = 0;
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_groups, groups.data(), groups.size() * sizeof(BGroup), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_jobs, bj.data(), bj.size() * sizeof(BJob), hipMemcpyHostToDevice, ctx->stream));
  static const unsigned long long bst_init[2] = {~0ull, ~0ull};
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_bstatus, bst_init, sizeof(bst_init), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_counts, 0, 2 * sizeof(unsigned long long), ctx->stream));

  float ms_geo = 0.f, ms_xfer = 0.f, t = 0.f;
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  /* pass 1: step counts of every ray; when every ray's slot of records
   * fits the budget, also its first slot records (most rays are then
   * complete and only the longer ones are continued, from their last slot
   * record), else a count alone and the records in pass 2 */
  /* the record budget: the set limit, or what the device can give the two
   * record arenas (free memory plus what they hold now, less 4 GiB for the
   * images, items and other buffers), whichever is smaller */
  size_t traj_bytes = ctx->batch_traj_bytes;
  {
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) == hipSuccess) {
      const size_t held = (ctx->traj_doubles + ctx->traj2_doubles) * sizeof(double);
      const size_t reserve = (size_t)4 << 30;
      const size_t avail = freeb + held > 2 * reserve ? freeb + held - reserve : (freeb + held) / 2;
      traj_bytes = std::min(traj_bytes, avail);
    }
  }
  long long budget = (long long)(traj_bytes / (REC * sizeof(double)));
  const int slot = ctx->batch_slot;
  bool try_slot = !turns && slot >= 2 && (long long)rays_total * slot <= budget;
  const bool pairs = !turns;   /* the single-lane trace counts subring turns */
  if ((rc = grow(ctx, ctx->d_nstep_ray, ctx->cap_nstep_ray, (size_t)rays_total))) return rc;
  if ((rc = grow(ctx, ctx->d_off, ctx->cap_off, (size_t)rays_total))) return rc;
  if (try_slot && (rc = grow(ctx, ctx->d_traj, ctx->traj_doubles, (size_t)rays_total * slot * REC))) {
    /* the device could not give the slots after all: count first, and hold
     * the passes to half of what failed */
    if (rc != IPGPU_ERR_NOMEM) return rc;
    try_slot = false;
    budget = std::min(budget, (long long)rays_total * slot / 2);
  }
  /* the first pass's tiles nearest their image's centre first: ray length
   * grows toward the photon ring near the centre, so the longest waves
   * start first and the pass ends on short ones instead of waiting on a
   * long wave dispatched last (modelled on the 64-job set: 56 -> 32 ms).
   * Each ray is traced the same way; only the order of the waves changes. */
  const long long ntiles = rays_total / WAVE;
  if (pairs) {
    std::vector<std::pair<float, int>> key((size_t)ntiles);
    for (int g = 0; g < ngroups; g++) {
      const int nx = groups[g].P.nx, ny = groups[g].P.ny, ty = (ny + TILE - 1) / TILE;
      const float inv = 1.f / (float)std::max(nx, ny);
      const long long t0 = groups[g].ray_base / WAVE;
      for (int t = 0; t < groups[g].nrays / WAVE; t++) {
        const int ti = t / ty, tj = t % ty;
        key[(size_t)(t0 + t)] = {std::hypot((float)(ti * TILE) + 0.5f * TILE - 0.5f * nx,
                                            (float)(tj * TILE) + 0.5f * TILE - 0.5f * ny) * inv,
                                 (int)(t0 + t)};
      }
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<float, int> &x, const std::pair<float, int> &y) { return x.first < y.first; });
    std::vector<int> wt((size_t)ntiles);
    for (long long t = 0; t < ntiles; t++) wt[(size_t)t] = key[(size_t)t].second;
    if ((rc = grow(ctx, ctx->d_wtile, ctx->cap_wtile, (size_t)ntiles))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_wtile, wt.data(), ntiles * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  }
  /* the records past the slot: chunks of an arena of ovf_arena_bytes, so
   * that the rays longer than the slot need no second trace (a ray that
   * cannot get all its chunks is continued as before) */
  double *arena = nullptr;
  unsigned long long nchunks = 0;
  int maxc = 0;
  if (pairs && try_slot && ctx->ovf_arena_bytes >= (size_t)OVF_CH * REC * sizeof(double)) {
    int mxn = 0;
    for (int g = 0; g < ngroups; g++) mxn = std::max(mxn, groups[g].P.maxnstep);
    maxc = std::min(64, std::max(1, (mxn - slot + OVF_CH) / OVF_CH));
    nchunks = ctx->ovf_arena_bytes / ((size_t)OVF_CH * REC * sizeof(double));
    if ((rc = grow(ctx, ctx->d_ovf_arena, ctx->ovf_arena_doubles, (size_t)nchunks * OVF_CH * REC))
        || (rc = grow(ctx, ctx->d_ovf_table, ctx->cap_ovf_table, (size_t)rays_total * maxc))
        || (rc = grow(ctx, ctx->d_ovf_ok, ctx->cap_ovf_ok, (size_t)rays_total))
        || (rc = grow(ctx, ctx->d_ovf_cnt, ctx->cap_ovf_cnt, (size_t)1))) {
      if (rc != IPGPU_ERR_NOMEM) return rc;
      nchunks = 0;   /* no arena: every long ray is continued */
    } else {
      HIPCHK(ctx, hipMemsetAsync(ctx->d_ovf_cnt, 0, sizeof(unsigned long long), ctx->stream));
      arena = ctx->d_ovf_arena;
    }
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  if (pairs)
    k_trace_slot<<<(unsigned)(2 * ntiles), WAVE, 0, ctx->stream>>>(ctx->geom, ctx->d_groups, ngroups, rays_total,
                                                                  try_slot ? slot : 0, ctx->d_traj, ctx->d_nstep_ray,
                                                                  ctx->d_wtile, arena, ctx->d_ovf_cnt, nchunks,
                                                                  ctx->d_ovf_table, maxc, ctx->d_ovf_ok);
  else
    k_trace_count<<<(unsigned)((rays_total + WAVE - 1) / WAVE), WAVE, 0, ctx->stream>>>(
        ctx->geom, ctx->d_groups, ngroups, rays_total, ctx->d_nstep_ray);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
  /* while the first pass runs: the rays with a pixel (ray_pixel's tiles
   * of TILE x TILE pixels, WAVE rays each), every ray's group and each
   * group's count of rays with a pixel */
  std::vector<char> valid(rays_total, 0);
  std::vector<int> &gid = ctx->h_gid;
  gid.resize(rays_total);
  std::vector<long long> gvalid(ngroups, 0);
  for (int g = 0; g < ngroups; g++) {
    const int nx = groups[g].P.nx, ny = groups[g].P.ny, tiles_y = (ny + TILE - 1) / TILE;
    const int tiles_x = (int)(groups[g].nrays / WAVE) / tiles_y;
    for (long long r = groups[g].ray_base; r < groups[g].ray_base + groups[g].nrays; r++) gid[r] = g;
    for (int tx = 0; tx < tiles_x; tx++)
      for (int ty = 0; ty < tiles_y; ty++) {
        const long long r0 = groups[g].ray_base + ((long long)tx * tiles_y + ty) * WAVE;
        for (int l = 0; l < WAVE; l++) {
          const bool in = tx * TILE + l / TILE < nx && ty * TILE + l % TILE < ny;
          valid[r0 + l] = in ? 1 : 0;
          gvalid[g] += in;
        }
      }
  }
  std::vector<int> ns_host(rays_total), ovf_host(arena ? rays_total : 0);
  HIPCHK(ctx, hipMemcpyAsync(ns_host.data(), ctx->d_nstep_ray, rays_total * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (arena)
    HIPCHK(ctx, hipMemcpyAsync(ovf_host.data(), ctx->d_ovf_ok, rays_total * sizeof(int), hipMemcpyDeviceToHost,
                               ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  hipEventElapsedTime(&t, ctx->ev[1], ctx->ev[2]);
  ms_geo += t;
  /* rays with a pixel, longest first (counting sort); the slot totals in
   * the same sequential pass */
  int mxs = 0;
  long long long_recs = 0, nlong = 0, slot_steps = 0;
  for (long long r = 0; r < rays_total; r++)
    if (valid[r]) {
      const int v = ns_host[r];
      mxs = v > mxs ? v : mxs;
      if (try_slot) {
        slot_steps += v;
        if (v >= slot) {
          long_recs += v + 1;
          nlong++;
        }
      }
    }
  std::vector<long long> bucket(mxs + 2, 0);
  for (long long r = 0; r < rays_total; r++)
    if (valid[r]) bucket[mxs - ns_host[r] + 1]++;
  for (int k = 1; k <= mxs + 1; k++) bucket[k] += bucket[k - 1];
  const long long nvalid = bucket[mxs + 1];
  long long nvalid_done = 0, steps_host = 0;
  std::vector<int> order(nvalid);
  for (long long r = 0; r < rays_total; r++)
    if (valid[r]) order[bucket[mxs - ns_host[r]]++] = (int)r;
  std::vector<long long> off_host(rays_total, 0);
  std::vector<int2> items;
  bool in_place = try_slot && (long long)rays_total * slot + long_recs <= budget;
  if (in_place && nlong > 0 && (rc = grow(ctx, ctx->d_traj2, ctx->traj2_doubles, (size_t)long_recs * REC))) {
    if (rc != IPGPU_ERR_NOMEM) return rc;
    in_place = false;   /* the passes below, in the memory the slots held */
    budget = std::min(budget, (long long)rays_total * slot);
  }
  if (in_place) {
    /* every ray's records in place: its slot, or for the nlong longest (at
     * the front of order) a full-length region of traj2 (off = -1 - offset) */
    std::vector<long long> off2(nlong > 0 ? nlong : 1);
    long long o2 = 0;
    for (long long r = 0; r < rays_total; r++) off_host[r] = (long long)r * slot;
    for (long long q = 0; q < nlong; q++) {
      const int r = order[q];
      off2[q] = o2;
      off_host[r] = -1 - o2;
      o2 += ns_host[r] + 1;
    }
    /* the long rays' continuation goes out first: the host builds the
     * transfer's items while it runs */
    if ((rc = grow(ctx, ctx->d_blk, ctx->cap_blk, (size_t)(nlong > 0 ? nlong : 1)))) return rc;
    if ((rc = grow(ctx, ctx->d_off2, ctx->cap_off2, (size_t)(nlong > 0 ? nlong : 1)))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_off, off_host.data(), rays_total * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    if (nlong > 0) {
      /* the long rays whose records past the slot are all in arena chunks
       * are copied to their regions; the others are continued from their
       * last slot record */
      std::vector<int> glist, rlist;
      std::vector<long long> goff, roff;
      for (long long q = 0; q < nlong; q++) {
        const int r = order[q];
        if (arena && ovf_host[r]) {
          glist.push_back(r);
          goff.push_back(off2[q]);
        } else {
          rlist.push_back(r);
          roff.push_back(off2[q]);
        }
      }
      if (!glist.empty()) {
        const int ng = (int)glist.size();
        if ((rc = grow(ctx, ctx->d_glist, ctx->cap_glist, (size_t)ng))
            || (rc = grow(ctx, ctx->d_goff2, ctx->cap_goff2, (size_t)ng)))
          return rc;
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_glist, glist.data(), ng * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_goff2, goff.data(), ng * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
        k_ovf_gather<<<(unsigned)ng, 256, 0, ctx->stream>>>(ctx->d_glist, ng, ctx->d_goff2, slot, ctx->d_nstep_ray,
                                                            ctx->d_traj, ctx->d_ovf_arena, ctx->d_ovf_table, maxc,
                                                            ctx->d_traj2);
        HIPCHK(ctx, hipGetLastError());
      }
      if (!rlist.empty()) {
        const int nrl = (int)rlist.size();
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_blk, rlist.data(), nrl * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipMemcpyAsync(ctx->d_off2, roff.data(), nrl * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
        k_slot_copy<<<1024, 256, 0, ctx->stream>>>(ctx->d_blk, nrl, ctx->d_off2, slot, ctx->d_traj, ctx->d_traj2);
        HIPCHK(ctx, hipGetLastError());
        k_trace_resume_list<<<(unsigned)((2 * nrl + WAVE - 1) / WAVE), WAVE, 0, ctx->stream>>>(
            ctx->geom, ctx->d_groups, ngroups, ctx->d_blk, nrl, ctx->d_off2, slot, ctx->d_nstep_ray, ctx->d_traj2,
            ctx->d_status);
        HIPCHK(ctx, hipGetLastError());
      }
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    /* the (job, ray) items, built in pinned host memory (the upload is then
     * a direct copy); every ray's group from a table of the groups' ranges */
    long long nit = 0;
    for (int g = 0; g < ngroups; g++)
      nit += gvalid[g] * (long long)(multi ? gsub[g].size() : members[g].size());
    if ((size_t)nit > ctx->h_items_cap) {
      if (ctx->h_items) hipHostFree(ctx->h_items);
      ctx->h_items = nullptr;
      ctx->h_items_cap = 0;
      HIPCHK(ctx, hipHostMalloc((void **)&ctx->h_items, (size_t)nit * sizeof(int2), hipHostMallocDefault));
      ctx->h_items_cap = (size_t)nit;
    }
    int2 *hit = ctx->h_items;
    long long w = 0;
    /* the rays by band, longest band first, and in ray-index order within a
     * band (a counting sort): a wave's items are then neighbouring rays of
     * one trace group where the band allows, whose records and dump cells
     * the caches share (the FP64 transfer 357 -> 345 ms against the rays'
     * step-count order within a band) */
    const int nband = (mxs >> ITEM_BAND) + 1;
    std::vector<long long> bstart(nband + 1, 0);
    for (long long r = 0; r < rays_total; r++)
      if (valid[r]) bstart[nband - (ns_host[r] >> ITEM_BAND)]++;
    for (int k = 1; k <= nband; k++) bstart[k] += bstart[k - 1];
    std::vector<int> border(nvalid);
    {
      std::vector<long long> fill(bstart.begin(), bstart.end() - 1);
      for (long long r = 0; r < rays_total; r++)
        if (valid[r]) border[fill[nband - 1 - (ns_host[r] >> ITEM_BAND)]++] = (int)r;
    }
    long long st_steps = 0, st_pairs = 0, st_psteps = 0;
    for (int k = 0; k < nband; k++)
      for (int sz = max_sz; sz >= 1; sz--)
        for (long long q = bstart[k]; q < bstart[k + 1]; q++) {
          const int r = border[q];
          const long long v = ns_host[r];
          ray_items(r, gid[r], sz, [&](int2 it) {
            hit[w++] = it;
            st_steps += v;
            st_pairs += sz;
            st_psteps += v * sz;
          });
        }
    const int nitems = (int)nit;
    ctx->bstat[0] += nit;
    ctx->bstat[1] += st_steps;
    ctx->bstat[3] += st_pairs;
    ctx->bstat[4] += st_psteps;
    if ((rc = grow(ctx, ctx->d_items, ctx->cap_items, (size_t)(nitems > 0 ? nitems : 1)))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_items, hit, nitems * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    if ((rc = batch_transfer(ctx, jobs[0].image.transfer_fp32 != 0, multi, nitems, nlong > 0 ? ctx->d_traj2 : nullptr,
                             nullptr, false)))
      return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    /* the records, items and job table stay resident: ipgpu_rerender_batch
     * can run this transfer again with new M_unit scales */
    ctx->blast.valid = true;
    ctx->blast.nitems = nitems;
    ctx->blast.traj2 = nlong > 0 ? ctx->d_traj2 : nullptr;
    ctx->blast.multi = multi;
    int mism = 0;
    HIPCHK(ctx, hipMemcpyAsync(&mism, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (mism) return set_err(ctx, IPGPU_ERR_HIP, "geodesic step counts differ between the two trace passes");
    hipEventElapsedTime(&t, ctx->ev[1], ctx->ev[2]);
    ms_geo += t;
    hipEventElapsedTime(&t, ctx->ev[4], ctx->ev[3]);
    ms_xfer += t;
    nvalid_done = nvalid;
    steps_host = slot_steps;
  }
  if (nvalid_done < nvalid && ctx->d_traj2) {
    /* passes use one record arena: the budget counted the second one's memory */
    HIPCHK(ctx, hipFree(ctx->d_traj2));
    ctx->dev_bytes -= ctx->traj2_doubles * sizeof(double);
    ctx->d_traj2 = nullptr;
    ctx->traj2_doubles = 0;
  }
  for (long long a = nvalid_done; a < nvalid;) {
    /* the longest prefix of the remaining rays whose records fit */
    long long b = a, recs = 0;
    while (b < nvalid && (b == a || recs + ns_host[order[b]] + 1 <= budget)) {
      off_host[order[b]] = recs;
      recs += ns_host[order[b]] + 1;
      b++;
    }
    const int nr = (int)(b - a);
    if ((rc = grow(ctx, ctx->d_traj, ctx->traj_doubles, (size_t)recs * REC))) {
      if (rc != IPGPU_ERR_NOMEM || b - a == 1) return rc;
      budget = recs / 2;   /* a smaller pass over the same rays */
      continue;
    }
    if (turns && (rc = grow(ctx, ctx->d_nturns, ctx->nturns_ints, (size_t)recs))) return rc;
    if ((rc = grow(ctx, ctx->d_blk, ctx->cap_blk, (size_t)nr))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_off, off_host.data(), rays_total * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_blk, order.data() + a, nr * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    if (pairs)
      k_trace_store_pair<<<(2 * nr + WAVE - 1) / WAVE, WAVE, 0, ctx->stream>>>(
          ctx->geom, ctx->d_groups, ngroups, ctx->d_blk, nr, ctx->d_off, ctx->d_nstep_ray, ctx->d_traj,
          ctx->d_status, ctx->d_counts);
    else
      k_trace_store<<<(nr + WAVE - 1) / WAVE, WAVE, 0, ctx->stream>>>(ctx->geom, ctx->d_groups, ngroups, ctx->d_blk,
                                                                     nr, ctx->d_off, ctx->d_nstep_ray, ctx->d_traj,
                                                                     turns ? ctx->d_nturns : nullptr, ctx->d_status,
                                                                     ctx->d_counts);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    /* work items (job, ray) of these rays, in the rays' (longest first) order */
    items.clear();
    for (long long q0 = a; q0 < b;) {
      long long q1 = q0;
      while (q1 < b && (ns_host[order[q1]] >> ITEM_BAND) == (ns_host[order[q0]] >> ITEM_BAND)) q1++;
      for (int sz = max_sz; sz >= 1; sz--)
        for (long long q = q0; q < q1; q++) {
          const int r = order[q];
          ray_items(r, group_of_host(groups, r), sz, [&](int2 v) { items.push_back(v); });
        }
      q0 = q1;
    }
    const int nitems = (int)items.size();
    for (const int2 &v : items) {
      const int r = multi ? v.x : v.y, nj = multi ? subsets[v.y].x : 1;
      ctx->bstat[0] += 1;
      ctx->bstat[1] += ns_host[r];
      ctx->bstat[3] += nj;
      ctx->bstat[4] += (long long)ns_host[r] * nj;
    }
    if ((rc = grow(ctx, ctx->d_items, ctx->cap_items, (size_t)(nitems > 0 ? nitems : 1)))) return rc;
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_items, items.data(), nitems * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    if ((rc = batch_transfer(ctx, jobs[0].image.transfer_fp32 != 0, multi, nitems, nullptr,
                             turns ? ctx->d_nturns : nullptr,
                             false)))
      return rc;
    HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    int mism = 0;
    HIPCHK(ctx, hipMemcpyAsync(&mism, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (mism) return set_err(ctx, IPGPU_ERR_HIP, "geodesic step counts differ between the two trace passes");
    hipEventElapsedTime(&t, ctx->ev[1], ctx->ev[2]);
    ms_geo += t;
    hipEventElapsedTime(&t, ctx->ev[4], ctx->ev[3]);
    ms_xfer += t;
    a = b;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  unsigned long long bst[2];
  HIPCHK(ctx, hipMemcpyAsync(bst, ctx->d_bstatus, sizeof(bst), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->counts, ctx->d_counts, sizeof(ctx->counts), hipMemcpyDeviceToHost, ctx->stream));
  unsigned long long bc = 0;
  HIPCHK(ctx, hipMemcpyAsync(&bc, ctx->d_bcounts, sizeof(bc), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->counts[0] += (unsigned long long)steps_host;   /* the rays whose records pass 1 stored */
  ctx->bstat[2] = multi ? (long long)bc : (long long)ctx->counts[1];
  hipEventElapsedTime(&ctx->ms_total, ctx->ev[0], ctx->ev[3]);
  ctx->ms_geo = ms_geo;
  ctx->ms_xfer = ms_xfer;
  ctx->ms_stage[0] = ms_geo;
  ctx->ms_stage[1] = 0.f;
  ctx->ms_stage[2] = ms_xfer;
  ctx->ms_stage[3] = 0.f;
  for (int q = 0; q < 2; q++)
    if (bst[q] != ~0ull) {
      unsigned long long key = bst[q] - 1;
      ctx->bad_job = (int)(key >> 32);
      int pix = (int)(key & 0xffffffffull);
      const ipgpu_image &P = jobs[ctx->bad_job].image;
      ctx->bad_i = pix / P.ny;
      ctx->bad_j = pix % P.ny;
      snprintf(ctx->err, sizeof(ctx->err), "%s in job %d pixel %d %d",
               q == 0 ? "MaxNStep exceeded" : "NaN in image", ctx->bad_job, ctx->bad_i, ctx->bad_j);
      return q == 0 ? IPGPU_ERR_MAXNSTEP : IPGPU_ERR_NAN;
    }
  if (ctx->blast.valid) {
    ctx->blast.njobs = njobs;
    ctx->blast.f32 = jobs[0].image.transfer_fp32 != 0;
    ctx->blast.img.resize(njobs);
    for (int j = 0; j < njobs; j++) ctx->blast.img[j] = jobs[j].image;
    ctx->blast.bj = bj;
    ctx->blast.pix_total = pix_total;
  }
  for (int j = 0; j < njobs; j++) {
    const ipgpu_job &J = jobs[j];
    size_t n = (size_t)J.image.nx * J.image.ny, o = (size_t)bj[j].pix_off;
    if (J.image_out) HIPCHK(ctx, hipMemcpyAsync(J.image_out, ctx->d_bimage + o, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (J.imageS_out && !J.image.only_unpolarized)
      HIPCHK(ctx, hipMemcpyAsync(J.imageS_out, ctx->d_bimageS + 5 * o, 5 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (J.taus_out) HIPCHK(ctx, hipMemcpyAsync(J.taus_out, ctx->d_btaus + o, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (J.nstep_out) HIPCHK(ctx, hipMemcpyAsync(J.nstep_out, ctx->d_bnstep + o, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (J.flag_out) HIPCHK(ctx, hipMemcpyAsync(J.flag_out, ctx->d_bflag + o, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IPGPU_OK;
}

/* The transfer of the resident batch as one stream sequence: the job table
 * from pinned memory, the status clears and the transfer launch. */
static int rerender_enqueue(ipgpu_ctx *ctx)
{
  const ipgpu_ctx::LastBatch &L = ctx->blast;
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_jobs, ctx->h_bjobs, (size_t)L.njobs * sizeof(BJob), hipMemcpyHostToDevice,
                             ctx->stream));
  return batch_transfer(ctx, L.f32 != 0, L.multi != 0, L.nitems, L.traj2, nullptr, true);
}

int ipgpu_rerender_batch(ipgpu_ctx *ctx, int njobs, const ipgpu_job *jobs)
{
  if (!ctx || njobs <= 0 || !jobs) return IPGPU_ERR_ARG;
  ipgpu_ctx::LastBatch &L = ctx->blast;
  if (!L.valid)
    return set_err(ctx, IPGPU_ERR_STATE, "no resident batch (none rendered, it ran in passes, or a render, "
                                         "adjoint or geometry change has reused its buffers since)");
  if (njobs != L.njobs) return set_err(ctx, IPGPU_ERR_ARG, "job count differs from the resident batch");
  for (int j = 0; j < njobs; j++) {
    if (!same_image(jobs[j].image, L.img[j]))
      return set_err(ctx, IPGPU_ERR_ARG, "a job's image differs from the resident batch's");
    if (!(jobs[j].m_unit_scale > 0.)) return set_err(ctx, IPGPU_ERR_ARG, "bad m_unit_scale");
  }
  if (!ctx->d_dump && ctx->model.kind == IPGPU_MODEL_IHARM)
    return set_err(ctx, IPGPU_ERR_STATE, "no fluid dump uploaded");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if ((size_t)njobs > ctx->h_bjobs_cap) {
    drop_batch_graph(ctx);   /* the graph reads the pinned table it was captured with */
    if (ctx->h_bjobs) hipHostFree(ctx->h_bjobs);
    ctx->h_bjobs = nullptr;
    ctx->h_bjobs_cap = 0;
    HIPCHK(ctx, hipHostMalloc((void **)&ctx->h_bjobs, (size_t)njobs * sizeof(BJob), hipHostMallocDefault));
    ctx->h_bjobs_cap = (size_t)njobs;
  }
  for (int j = 0; j < njobs; j++) {
    BJob b = L.bj[j];
    b.sc.ne = jobs[j].m_unit_scale;
    b.sc.b = sqrt(jobs[j].m_unit_scale);
    ctx->h_bjobs[j] = b;
  }
  if (L.f32) {
    int rc0 = ensure_dump32(ctx);   /* before any capture: it may allocate */
    if (rc0) return rc0;
  }
  const int nw = ctx->batch_queue_waves > 0 ? ctx->batch_queue_waves
                                            : (L.multi ? ctx->xm_occ : ctx->xq_occ)[L.f32 ? 1 : 0] * ctx->ncu;
  ipgpu_ctx::GraphKey key;
  memset(&key, 0, sizeof(key));
  key.geom = ctx->geom;
  key.model = ctx->model;
  key.dump = ctx->d_dump;
  key.dump32 = L.f32 ? ctx->d_dump32 : nullptr;
  key.queue = ctx->batch_queue;
  key.waves = nw;
  if (ctx->bgraph_exec && memcmp(&key, &ctx->bkey, sizeof(key)) != 0) drop_batch_graph(ctx);
  HIPCHK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  int rc;
  if (ctx->rerender_graph) {
    if (!ctx->bgraph_exec) {
      HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      rc = rerender_enqueue(ctx);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
      }
      HIPCHK(ctx, e);
      ctx->bgraph = g;
      HIPCHK(ctx, hipGraphInstantiate(&ctx->bgraph_exec, g, nullptr, nullptr, 0));
      ctx->bkey = key;
    }
    HIPCHK(ctx, hipGraphLaunch(ctx->bgraph_exec, ctx->stream));
  } else if ((rc = rerender_enqueue(ctx))) {
    return rc;
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  unsigned long long bst[2];
  HIPCHK(ctx, hipMemcpyAsync(bst, ctx->d_bstatus, sizeof(bst), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(ctx->counts, ctx->d_counts, sizeof(ctx->counts), hipMemcpyDeviceToHost, ctx->stream));
  unsigned long long bc = 0;
  HIPCHK(ctx, hipMemcpyAsync(&bc, ctx->d_bcounts, sizeof(bc), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->counts[0] = 0;   /* no geodesic is traced */
  ctx->bstat[2] = L.multi ? (long long)bc : (long long)ctx->counts[1];
  hipEventElapsedTime(&ctx->ms_total, ctx->ev[0], ctx->ev[3]);
  ctx->ms_geo = 0.f;
  ctx->ms_xfer = ctx->ms_total;
  ctx->ms_stage[0] = 0.f;
  ctx->ms_stage[1] = 0.f;
  ctx->ms_stage[2] = ctx->ms_total;
  ctx->ms_stage[3] = 0.f;
  for (int q = 0; q < 2; q++)
    if (bst[q] != ~0ull) {
      unsigned long long k = bst[q] - 1;
      ctx->bad_job = (int)(k >> 32);
      int pix = (int)(k & 0xffffffffull);
      const ipgpu_image &P = jobs[ctx->bad_job].image;
      ctx->bad_i = pix / P.ny;
      ctx->bad_j = pix % P.ny;
      snprintf(ctx->err, sizeof(ctx->err), "%s in job %d pixel %d %d",
               q == 0 ? "MaxNStep exceeded" : "NaN in image", ctx->bad_job, ctx->bad_i, ctx->bad_j);
      return q == 0 ? IPGPU_ERR_MAXNSTEP : IPGPU_ERR_NAN;
    }
  for (int j = 0; j < njobs; j++) {
    const ipgpu_job &J = jobs[j];
    size_t n = (size_t)J.image.nx * J.image.ny, o = (size_t)L.bj[j].pix_off;
    if (J.image_out) HIPCHK(ctx, hipMemcpyAsync(J.image_out, ctx->d_bimage + o, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (J.imageS_out && !J.image.only_unpolarized)
      HIPCHK(ctx, hipMemcpyAsync(J.imageS_out, ctx->d_bimageS + 5 * o, 5 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (J.taus_out) HIPCHK(ctx, hipMemcpyAsync(J.taus_out, ctx->d_btaus + o, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (J.nstep_out) HIPCHK(ctx, hipMemcpyAsync(J.nstep_out, ctx->d_bnstep + o, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (J.flag_out) HIPCHK(ctx, hipMemcpyAsync(J.flag_out, ctx->d_bflag + o, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IPGPU_OK;
}

int ipgpu_set_batch_options(ipgpu_ctx *ctx, int slot, int queue, int queue_waves, int rerender_graph)
{
  if (!ctx) return IPGPU_ERR_ARG;
  if (slot > 0) {
    if (slot < 2) return set_err(ctx, IPGPU_ERR_ARG, "record slot below 2");
    ctx->batch_slot = slot;
  }
  if (queue >= 0) ctx->batch_queue = queue != 0;
  if (queue_waves >= 0) ctx->batch_queue_waves = queue_waves;
  if (rerender_graph >= 0) ctx->rerender_graph = rerender_graph != 0;
  drop_batch_graph(ctx);
  return IPGPU_OK;
}

int ipgpu_autotune_batch(ipgpu_ctx *ctx, int njobs, const ipgpu_job *jobs, int chosen[4])
{
  if (!ctx || njobs <= 0 || !jobs) return IPGPU_ERR_ARG;
  std::vector<ipgpu_job> q(jobs, jobs + njobs);
  for (ipgpu_job &j : q) {
    j.image_out = j.imageS_out = j.taus_out = nullptr;
    j.nstep_out = j.flag_out = nullptr;
  }
  const int slot0 = ctx->batch_slot, queue0 = ctx->batch_queue, waves0 = ctx->batch_queue_waves;
  const int graph0 = ctx->rerender_graph;
  auto restore = [&](int rc) {
    ctx->batch_slot = slot0;
    ctx->batch_queue = queue0;
    ctx->batch_queue_waves = waves0;
    ctx->rerender_graph = graph0;
    return rc;
  };
  /* 1. the record slot of the first trace: whole batches, device time.
   * Largest slot first, after an untimed batch at it: the record arena is
   * then grown before any timed batch (an arena grown inside a timed batch
   * made its slot lose by the allocation's cost) */
  static const int slots[3] = {4096, 2048, 1024};
  int best_slot = slot0, rc;
  float best = 3.4e38f;
  ctx->batch_slot = slots[0];
  if ((rc = ipgpu_render_batch(ctx, njobs, q.data()))) return restore(rc);
  for (int s : slots) {
    ctx->batch_slot = s;
    if ((rc = ipgpu_render_batch(ctx, njobs, q.data()))) return restore(rc);
    /* the trace and transfer launches (the first use of a larger record
     * arena also pays for its allocation, which later batches do not) */
    if (ctx->ms_geo + ctx->ms_xfer < best) {
      best = ctx->ms_geo + ctx->ms_xfer;
      best_slot = s;
    }
  }
  ctx->batch_slot = best_slot;
  if (best_slot != slots[2] && (rc = ipgpu_render_batch(ctx, njobs, q.data()))) return restore(rc);
  /* 2. the transfer launch: one wave per 64 items, or the work queue at one
   * and two times the resident waves; the transfer alone, on the records
   * left in place (a batch in passes keeps the default launch) */
  int best_q = queue0, best_w = waves0;
  if (ctx->blast.valid) {
    const int occ = ctx->xq_occ[ctx->blast.f32 ? 1 : 0] * ctx->ncu;
    const int cand[3][2] = {{0, 0}, {1, occ}, {1, 2 * occ}};
    ctx->rerender_graph = 0;
    best = 3.4e38f;
    for (const auto &c : cand) {
      ctx->batch_queue = c[0];
      ctx->batch_queue_waves = c[1];
      if ((rc = ipgpu_rerender_batch(ctx, njobs, q.data()))) return restore(rc);
      if (ctx->ms_total < best) {
        best = ctx->ms_total;
        best_q = c[0];
        best_w = c[1];
      }
    }
    ctx->rerender_graph = graph0;
  }
  ctx->batch_queue = best_q;
  ctx->batch_queue_waves = best_w;
  drop_batch_graph(ctx);
  if (chosen) {
    chosen[0] = ctx->batch_slot;
    chosen[1] = ctx->batch_queue;
    chosen[2] = ctx->batch_queue > 0 ? (ctx->batch_queue_waves > 0 ? ctx->batch_queue_waves
                                                                   : ctx->xq_occ[q[0].image.transfer_fp32 ? 1 : 0] * ctx->ncu)
                                     : 0;
    chosen[3] = ctx->xq_occ[q[0].image.transfer_fp32 ? 1 : 0];
  }
  return IPGPU_OK;
}

int ipgpu_batch_counts(const ipgpu_ctx *ctx, long long out[5])
{
  if (!ctx || !out) return IPGPU_ERR_ARG;
  for (int k = 0; k < 5; k++) out[k] = ctx->bstat[k];
  return IPGPU_OK;
}

int ipgpu_set_render_memory(ipgpu_ctx *ctx, size_t bytes)
{
  if (!ctx || bytes < ((size_t)1 << 20)) return set_err(ctx, IPGPU_ERR_ARG, "render memory below 1 MiB");
  ctx->render_bytes = bytes;
  return IPGPU_OK;
}

int ipgpu_set_gather_staging(ipgpu_ctx *ctx, int on)
{
  if (!ctx) return IPGPU_ERR_ARG;
  ctx->gather_stage = on ? 1 : 0;
  return IPGPU_OK;
}

int ipgpu_render_info(const ipgpu_ctx *ctx, int *chunks, int *continued_rays, size_t *device_bytes)
{
  if (!ctx) return IPGPU_ERR_ARG;
  if (chunks) *chunks = ctx->last_chunks;
  if (continued_rays) *continued_rays = ctx->last_overflow;
  if (device_bytes) *device_bytes = ctx->dev_bytes;
  return IPGPU_OK;
}

int ipgpu_set_batch_memory(ipgpu_ctx *ctx, size_t traj_bytes)
{
  if (!ctx || traj_bytes < ((size_t)1 << 20)) return IPGPU_ERR_ARG;
  ctx->batch_traj_bytes = traj_bytes;
  return IPGPU_OK;
}

int ipgpu_bad_job(const ipgpu_ctx *ctx, int *job, int *i, int *j)
{
  if (!ctx || !job || !i || !j) return IPGPU_ERR_ARG;
  *job = ctx->bad_job;
  *i = ctx->bad_i;
  *j = ctx->bad_j;
  return IPGPU_OK;
}

/* ------------------------------------------------------------------ adjoint */

int ipgpu_scale_density(ipgpu_ctx *ctx, const double *scale)
{
  if (!ctx || !scale) return IPGPU_ERR_ARG;
  if (!ctx->d_dump || ctx->model.kind != IPGPU_MODEL_IHARM) return set_err(ctx, IPGPU_ERR_STATE, "no iharm dump");
  if (ctx->geom.N3 % 2) return set_err(ctx, IPGPU_ERR_ARG, "per-cell density scale needs an even N3");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const ipgpu_geom &G = ctx->geom;
  const size_t nin = (size_t)G.N1 * G.N2 * G.N3;
  int rc;
  if ((rc = grow(ctx, ctx->d_scale, ctx->cap_scale, nin))) return rc;
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_scale, scale, nin * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  k_scale_ne<<<(unsigned)((nin + 255) / 256), 256, 0, ctx->stream>>>(G, ctx->d_scale, ctx->d_dump);
  int nr = G.N2 * G.N3, np = (G.N1 + 2) * G.N3, nf = (G.N1 + 2) * (G.N2 + 2);
  k_bc_ne<<<(nr + 255) / 256, 256, 0, ctx->stream>>>(G, 0, ctx->d_dump);
  k_bc_ne<<<(np + 255) / 256, 256, 0, ctx->stream>>>(G, 1, ctx->d_dump);
  k_bc_ne<<<(nf + 255) / 256, 256, 0, ctx->stream>>>(G, 2, ctx->d_dump);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return IPGPU_OK;
}

/* Fold the gradient of ghost cells back onto the interior cells they copy
 * (the transpose of the ghost-zone rules, applied in reverse order:
 * azimuth, polar, radial), then extract the interior.  Even N3 (the polar
 * rule is then a permutation, kflip). */
static void fold_ghosts(const ipgpu_geom &G, std::vector<double> &g, double *interior)
{
  const int N1 = G.N1, N2 = G.N2, N3 = G.N3;
  auto at = [&](int i, int j, int k) -> double & { return g[((size_t)i * (N2 + 2) + j) * (N3 + 2) + k]; };
  for (int i = 0; i < N1 + 2; i++)
    for (int j = 0; j < N2 + 2; j++) {
      at(i, j, N3) += at(i, j, 0);
      at(i, j, 1) += at(i, j, N3 + 1);
      at(i, j, 0) = at(i, j, N3 + 1) = 0.;
    }
  for (int i = 0; i < N1 + 2; i++)
    for (int k = 1; k <= N3; k++) {
      int kflip = ((k - 1) + (N3 / 2)) % N3 + 1;
      at(i, 1, kflip) += at(i, 0, k);
      at(i, N2, kflip) += at(i, N2 + 1, k);
      at(i, 0, k) = at(i, N2 + 1, k) = 0.;
    }
  for (int j = 1; j <= N2; j++)
    for (int k = 1; k <= N3; k++) {
      at(1, j, k) += at(0, j, k);
      at(N1, j, k) += at(N1 + 1, j, k);
    }
  for (int i = 0; i < N1; i++)
    for (int j = 0; j < N2; j++)
      for (int k = 0; k < N3; k++) interior[((size_t)i * N2 + j) * N3 + k] = at(i + 1, j + 1, k + 1);
}

/* The adjoint pass shared by ipgpu_adjoint (target and weights: the L2
 * loss) and ipgpu_backward (cot: the gradient of any loss with respect to
 * the output image). */
static int run_adjoint(ipgpu_ctx *ctx, const ipgpu_image *img, const double *target, const double *weights,
                       const double *cot, double out4[4], double *grad_density, double *imageS)
{
  if (!ctx || !img || !(target || cot) || !out4) return IPGPU_ERR_ARG;
  if (!ctx->has_geom || !ctx->has_model || !ctx->d_dump || ctx->model.kind != IPGPU_MODEL_IHARM)
    return set_err(ctx, IPGPU_ERR_STATE, "the adjoint needs an iharm dump");
  if (!ctx->has_units) return set_err(ctx, IPGPU_ERR_STATE, "the adjoint needs a dump from ipgpu_upload_prims");
  batch_evict(ctx);
  release_overflow(ctx);
  const ipgpu_image &P = *img;
  if (P.nx <= 0 || P.ny <= 0 || P.maxnstep < 2) return set_err(ctx, IPGPU_ERR_ARG, "bad image size or maxnstep");
  if (P.only_unpolarized || P.transfer_fp32 || P.target_nturns >= 0 || P.counterjet)
    return set_err(ctx, IPGPU_ERR_ARG, "the adjoint covers the polarized FP64 path without subring or counterjet selection");
  if (P.emission_type != 4 && P.emission_type != 1) return set_err(ctx, IPGPU_ERR_ARG, "emission_type must be 4 or 1");
  if (grad_density && ctx->geom.N3 % 2) return set_err(ctx, IPGPU_ERR_ARG, "per-cell density gradient needs an even N3");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const int tiles = ((P.nx + TILE - 1) / TILE) * ((P.ny + TILE - 1) / TILE);
  const int nrays = tiles * WAVE;
  const size_t npix = (size_t)P.nx * P.ny;
  const ipgpu_geom &G = ctx->geom;
  const size_t nc = (size_t)(G.N1 + 2) * (G.N2 + 2) * (G.N3 + 2);
  int rc;
  if ((rc = grow(ctx, ctx->d_nstep_ray, ctx->cap_nstep_ray, (size_t)nrays))) return rc;
  if ((rc = grow(ctx, ctx->d_imageS, ctx->cap_imageS, npix * 5))) return rc;
  if ((rc = grow(ctx, ctx->d_target, ctx->cap_target, npix * 5))) return rc;
  if ((rc = grow(ctx, ctx->d_out4, ctx->cap_out4, (size_t)nrays))) return rc;
  if ((rc = grow(ctx, ctx->d_gcell, ctx->cap_gcell, nc))) return rc;
  int cap, stride;
  unsigned long long total;
  if ((rc = prepare_packets(ctx, P, nrays, true, cap, total, stride))) return rc;
  /* segments of every ray (prepare_packets has synchronised: the step
   * counts are final) */
  std::vector<int> &ns = ctx->h_ns;
  ns.resize(nrays);
  HIPCHK(ctx, hipMemcpy(ns.data(), ctx->d_nstep_ray, nrays * sizeof(int), hipMemcpyDeviceToHost));
  const int Ls = ctx->adj_seg_len;
  std::vector<int> &sr = ctx->h_aseg, &lanes = ctx->h_alane, &r0 = ctx->h_aray;
  sr.clear();
  lanes.clear();
  r0.assign((size_t)nrays + 1, 0);
  for (int r = 0; r < nrays; r++) {
    r0[r] = (int)(sr.size() / 3);
    if (ns[r] <= 0 || ns[r] >= P.maxnstep - 1) continue;
    for (int q0 = 0, k = 0; q0 < ns[r]; q0 += Ls, k++) {
      const int sg = (int)(sr.size() / 3);
      sr.push_back(r);
      sr.push_back(q0);
      sr.push_back(q0 + Ls < ns[r] ? q0 + Ls : ns[r]);
      for (int b = (k == 0 ? NPAR : 0); b <= NPAR; b++) {
        lanes.push_back(sg);
        lanes.push_back(b);
      }
    }
  }
  const int nseg = (int)(sr.size() / 3);
  r0[nrays] = nseg;
  const int nlanes = (int)(lanes.size() / 2);
  const int nck = (Ls + CK - 1) / CK;
  /* segment table: ray, q0, q1 rows; lanes: segment, column rows; ray
   * starts */
  if ((rc = grow(ctx, ctx->d_aint, ctx->cap_aint, (size_t)3 * nseg + 2 * nlanes + nrays + 1 + 3))) return rc;
  int *d_sray = ctx->d_aint, *d_sq0 = d_sray + nseg, *d_sq1 = d_sq0 + nseg;
  int *d_lseg = d_sq1 + nseg, *d_lcol = d_lseg + nlanes, *d_r0 = d_lcol + nlanes;
  {
    std::vector<int> &t = ctx->h_tmp;
    t.resize((size_t)3 * nseg + 2 * nlanes + nrays + 1);
    for (int k = 0; k < nseg; k++) {
      t[k] = sr[3 * k];
      t[nseg + k] = sr[3 * k + 1];
      t[2 * nseg + k] = sr[3 * k + 2];
    }
    for (int l = 0; l < nlanes; l++) {
      t[3 * (size_t)nseg + l] = lanes[2 * l];
      t[3 * (size_t)nseg + nlanes + l] = lanes[2 * l + 1];
    }
    for (int r = 0; r <= nrays; r++) t[3 * (size_t)nseg + 2 * nlanes + r] = r0[r];
    HIPCHK(ctx, hipMemcpyAsync(ctx->d_aint, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  }
  const size_t nsg = nseg > 0 ? (size_t)nseg : 1;
  if ((rc = grow(ctx, ctx->d_amap, ctx->cap_amap, nsg * ASEG_MAP)) || (rc = grow(ctx, ctx->d_aN, ctx->cap_aN, nsg))
      || (rc = grow(ctx, ctx->d_aL, ctx->cap_aL, nsg)) || (rc = grow(ctx, ctx->d_ckpt, ctx->cap_ckpt, nsg * nck))
      || (rc = grow(ctx, ctx->d_seg, ctx->cap_seg, nsg * CK)) || (rc = grow(ctx, ctx->d_ag3, ctx->cap_ag3, nsg * 3)))
    return rc;
  HIPCHK(ctx, hipMemcpyAsync(ctx->d_target, cot ? cot : target, npix * 5 * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_gcell, 0, nc * sizeof(double), ctx->stream));
  AdjW w;
  for (int q = 0; q < 4; q++) w.w[q] = weights ? weights[q] : 1.;
  if (nlanes > 0) {
    k_adj_map<<<(nlanes + WAVE - 1) / WAVE, WAVE, 0, ctx->stream>>>(nlanes, d_lseg, d_lcol, d_sray, d_sq0, d_sq1,
                                                                    ctx->d_off, ctx->d_pkt, ctx->d_amap);
    HIPCHK(ctx, hipGetLastError());
  }
  k_adj_compose<<<nrays / WAVE, WAVE, 0, ctx->stream>>>(G, P, nrays, ctx->d_nstep_ray, d_r0, ctx->d_amap,
                                                         cot ? nullptr : ctx->d_target, w,
                                                         cot ? ctx->d_target : nullptr, ctx->d_aN, ctx->d_aL,
                                                         ctx->d_out4, ctx->d_imageS, ctx->d_status);
  HIPCHK(ctx, hipGetLastError());
  if (nseg > 0) {
    k_adj_seg<<<(nseg + WAVE - 1) / WAVE, WAVE, 0, ctx->stream>>>(G, P, nseg, d_sray, d_sq0, d_sq1, ctx->d_off,
                                                                  ctx->d_pkt, ctx->d_aux, ctx->d_dump, ctx->d_aN,
                                                                  ctx->d_aL, nck, ctx->d_ckpt, ctx->d_seg,
                                                                  ctx->d_gcell, ctx->d_ag3);
    HIPCHK(ctx, hipGetLastError());
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
  std::vector<double> lossv((size_t)nrays), g3((size_t)nseg * 3);
  int status[NSTATUS];
  HIPCHK(ctx, hipMemcpyAsync(lossv.data(), ctx->d_out4, lossv.size() * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  if (nseg > 0)
    HIPCHK(ctx, hipMemcpyAsync(g3.data(), ctx->d_ag3, g3.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(status, ctx->d_status, sizeof(status), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  hipEventElapsedTime(&ctx->ms_total, ctx->ev[0], ctx->ev[3]);
  hipEventElapsedTime(&ctx->ms_geo, ctx->ev[0], ctx->ev[1]);
  hipEventElapsedTime(&ctx->ms_xfer, ctx->ev[1], ctx->ev[3]);
  if (status[1] != 0x7fffffff) {
    int p = status[1] - 1;
    ctx->bad_i = p / P.ny;
    ctx->bad_j = p % P.ny;
    snprintf(ctx->err, sizeof(ctx->err), "MaxNStep exceeded in pixel %d %d", ctx->bad_i, ctx->bad_j);
    return IPGPU_ERR_MAXNSTEP;
  }
  double acc[4] = {0., 0., 0., 0.};
  for (int r = 0; r < nrays; r++) acc[0] += lossv[r];
  for (int k = 0; k < nseg; k++)
    for (int q = 0; q < 3; q++) acc[q + 1] += g3[(size_t)k * 3 + q];
  out4[0] = acc[0];
  out4[1] = acc[1] / ctx->units.M_unit;
  out4[2] = acc[2];
  out4[3] = acc[3];
  if (grad_density) {
    std::vector<double> g(nc);
    HIPCHK(ctx, hipMemcpy(g.data(), ctx->d_gcell, nc * sizeof(double), hipMemcpyDeviceToHost));
    fold_ghosts(G, g, grad_density);
  }
  if (imageS) HIPCHK(ctx, hipMemcpy(imageS, ctx->d_imageS, npix * 5 * sizeof(double), hipMemcpyDeviceToHost));
  return IPGPU_OK;
}

int ipgpu_adjoint(ipgpu_ctx *ctx, const ipgpu_image *img, const double *target, const double *weights,
                  double out4[4], double *grad_density, double *imageS)
{
  if (!target) return set_err(ctx, IPGPU_ERR_ARG, "ipgpu_adjoint: no target image");
  return run_adjoint(ctx, img, target, weights, nullptr, out4, grad_density, imageS);
}

int ipgpu_backward(ipgpu_ctx *ctx, const ipgpu_image *img, const double *dL_dimageS, double out3[3],
                   double *grad_density, double *imageS)
{
  if (!dL_dimageS || !out3) return set_err(ctx, IPGPU_ERR_ARG, "ipgpu_backward: no output gradient");
  double out4[4];
  int rc = run_adjoint(ctx, img, nullptr, nullptr, dL_dimageS, out4, grad_density, imageS);
  if (rc) return rc;
  for (int q = 0; q < 3; q++) out3[q] = out4[q + 1];
  return IPGPU_OK;
}

int ipgpu_step_counts(const ipgpu_ctx *ctx, long long *geodesic_steps, long long *radiating_steps)
{
  if (!ctx) return IPGPU_ERR_ARG;
  if (geodesic_steps) *geodesic_steps = (long long)ctx->counts[0];
  if (radiating_steps) *radiating_steps = (long long)ctx->counts[1];
  return IPGPU_OK;
}

int ipgpu_bad_pixel(const ipgpu_ctx *ctx, int *i, int *j)
{
  if (!ctx || !i || !j) return IPGPU_ERR_ARG;
  *i = ctx->bad_i;
  *j = ctx->bad_j;
  return IPGPU_OK;
}

int ipgpu_kernel_times(const ipgpu_ctx *ctx, double *g, double *x, double *t)
{
  if (!ctx) return IPGPU_ERR_ARG;
  if (g) *g = ctx->ms_geo;
  if (x) *x = ctx->ms_xfer;
  if (t) *t = ctx->ms_total;
  return IPGPU_OK;
}

int ipgpu_trace_pixel(ipgpu_ctx *ctx, const ipgpu_image *img, int i, int j, int cap, int *nstep, double *out)
{
  if (!ctx || !img || !nstep || !out) return IPGPU_ERR_ARG;
  if (!ctx->has_geom || !ctx->has_model) return set_err(ctx, IPGPU_ERR_STATE, "geometry/model not set");
  if (ctx->model.kind == IPGPU_MODEL_IHARM && !ctx->d_dump)
    return set_err(ctx, IPGPU_ERR_STATE, "no fluid dump uploaded");
  const ipgpu_image &P = *img;
  if (i < 0 || j < 0 || i >= P.nx || j >= P.ny) return set_err(ctx, IPGPU_ERR_ARG, "pixel outside the image");
  if (P.maxnstep < 2 || cap < P.maxnstep) return set_err(ctx, IPGPU_ERR_ARG, "trace capacity below maxnstep");
  batch_evict(ctx);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  int rc;
  if ((rc = grow(ctx, ctx->d_trrec, ctx->cap_trrec, (size_t)P.maxnstep * REC))) return rc;
  if ((rc = grow(ctx, ctx->d_trout, ctx->cap_trout, (size_t)P.maxnstep * IPGPU_TRACE_NF))) return rc;
  if ((rc = grow(ctx, ctx->d_trnt, ctx->cap_trnt, (size_t)P.maxnstep))) return rc;
  HIPCHK(ctx, hipMemsetAsync(ctx->d_trout, 0, (size_t)P.maxnstep * IPGPU_TRACE_NF * sizeof(double), ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
  k_trace_pixel<<<1, WAVE, 0, ctx->stream>>>(ctx->geom, ctx->model, P, i, j, P.maxnstep, ctx->has_units ? 1 : 0,
                                              ctx->d_dump, ctx->d_trrec, ctx->d_trnt, ctx->d_trout, ctx->d_status);
  HIPCHK(ctx, hipGetLastError());
  int ns = 0;
  HIPCHK(ctx, hipMemcpyAsync(&ns, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *nstep = ns;
  if (ns >= P.maxnstep - 1) return set_err(ctx, IPGPU_ERR_MAXNSTEP, "MaxNStep exceeded in the traced pixel");
  HIPCHK(ctx, hipMemcpy(out, ctx->d_trout, (size_t)P.maxnstep * IPGPU_TRACE_NF * sizeof(double),
                        hipMemcpyDeviceToHost));
  return IPGPU_OK;
}

int ipgpu_stage_times(const ipgpu_ctx *ctx, double ms[4])
{
  if (!ctx || !ms) return IPGPU_ERR_ARG;
  for (int q = 0; q < 4; q++) ms[q] = ctx->ms_stage[q];
  return ctx->streamed;
}

int ipgpu_set_tuning(ipgpu_ctx *ctx, int trace_budget, int trace_phases, int seg_min_steps, int seg_len,
                     int record_steps)
{
  if (!ctx) return IPGPU_ERR_ARG;
  if (trace_budget >= 0) ctx->trace_budget = trace_budget;
  if (trace_phases >= 0) ctx->trace_phases = trace_phases;
  if (seg_min_steps >= 0) ctx->seg_min_steps = seg_min_steps;
  if (seg_len >= 0) ctx->seg_len = seg_len > 1 ? seg_len : 1;
  if (record_steps >= 0) ctx->traj_cap_steps = record_steps > 2 ? record_steps : 2;
  return IPGPU_OK;
}

int ipgpu_autotune_render(ipgpu_ctx *ctx, const ipgpu_image *img, int chosen[4])
{
  if (!ctx || !img) return IPGPU_ERR_ARG;
  const int tb0 = ctx->trace_budget, ph0 = ctx->trace_phases, sm0 = ctx->seg_min_steps, sl0 = ctx->seg_len;
  /* device time of one setting: the better of two renders after a first
   * one that grows the arenas for it */
  auto timed = [&](float &ms) -> int {
    float best = 3.4e38f;
    for (int rep = 0; rep < 3; rep++) {
      int rc = ipgpu_render(ctx, img, nullptr, nullptr, nullptr, nullptr, nullptr);
      if (rc) return rc;
      if (rep > 0 && ctx->ms_total < best) best = ctx->ms_total;
    }
    ms = best;
    return 0;
  };
  auto fail = [&](int rc) {
    ctx->trace_budget = tb0;
    ctx->trace_phases = ph0;
    ctx->seg_min_steps = sm0;
    ctx->seg_len = sl0;
    return rc;
  };
  int rc;
  float best;
  if ((rc = timed(best))) return fail(rc);
  /* one setting at a time, the others held at the best so far: segment
   * length, the segment threshold (every ray of the last phase, or also
   * those of the phase before), then the budget of the bounded phases */
  const int seg_lens[2] = {256, 512};
  for (int v : seg_lens) {
    const int keep = ctx->seg_len;
    ctx->seg_len = v;
    float ms;
    if ((rc = timed(ms))) return fail(rc);
    if (ms < best) best = ms;
    else ctx->seg_len = keep;
  }
  {
    const int keep = ctx->seg_min_steps;
    const int last = ctx->trace_budget, nb = ctx->trace_phases > 1 ? ctx->trace_phases : 1;
    const int cand[2] = {last, (int)((long long)last * (nb > 1 ? nb - 1 : 1) / (nb > 1 ? nb : 2))};
    for (int v : cand) {
      if (v == keep) continue;
      ctx->seg_min_steps = v;
      float ms;
      if ((rc = timed(ms))) return fail(rc);
      if (ms < best) best = ms;
      else ctx->seg_min_steps = keep;
    }
  }
  {
    const int keep_b = ctx->trace_budget, keep_m = ctx->seg_min_steps;
    const int cand[2] = {keep_b * 3 / 4, keep_b * 5 / 4};
    for (int v : cand) {
      if (v < 256) continue;
      ctx->trace_budget = v;
      /* a threshold at the old budget follows it */
      if (keep_m == keep_b) ctx->seg_min_steps = v;
      float ms;
      if ((rc = timed(ms))) return fail(rc);
      if (ms < best) {
        best = ms;
        break;
      }
      ctx->trace_budget = keep_b;
      ctx->seg_min_steps = keep_m;
    }
  }
  if (chosen) {
    chosen[0] = ctx->trace_budget;
    chosen[1] = ctx->trace_phases;
    chosen[2] = ctx->seg_min_steps;
    chosen[3] = ctx->seg_len;
  }
  ctx->ms_total = best;
  return IPGPU_OK;
}

/* native_coord + root_find, src/coordinates.c:375-442 */
int ipgpu_native_coord(const ipgpu_geom *G, double r, double thd, double phid, double X[4])
{
  if (!G || !X) return IPGPU_ERR_ARG;
  if (G->metric == IPGPU_METRIC_EMINKOWSKI) {
    X[0] = 1;
    X[1] = log(r);
    X[2] = thd / 180 * M_PI;
    X[3] = phid / 180 * M_PI;
    return IPGPU_OK;
  }
  double th = thd / 180. * M_PI;
  double x1 = log(r), xa, xb, xc;
  if (th < M_PI / 2.) {
    xa = G->cstartx[2];
    xb = (G->cstopx[2] - G->cstartx[2]) / 2 + SMALL;
  } else {
    xa = (G->cstopx[2] - G->cstartx[2]) / 2 - SMALL;
    xb = G->cstopx[2];
  }
  auto theta = [&](double x2) {
    double rr, t;
    bl_coord(*G, x1, x2, rr, t);
    return t;
  };
  double tol = 1.e-9, tha = theta(xa), thb = theta(xb), thc;
  X[0] = 0.0;
  X[1] = x1;
  X[3] = phid / 180. * M_PI;
  if (fabs(tha - th) < tol) {
    X[2] = xa;
    return IPGPU_OK;
  } else if (fabs(thb - th) < tol) {
    X[2] = xb;
    return IPGPU_OK;
  }
  xc = xa;
  for (int i = 0; i < 1000; i++) {
    xc = 0.5 * (xa + xb);
    thc = theta(xc);
    if ((thc - th) * (thb - th) < 0.) xa = xc;
    else xb = xc;
    if (fabs(theta(xc) - th) < tol) break;
  }
  X[2] = xc;
  return IPGPU_OK;
}

}  // extern "C"
