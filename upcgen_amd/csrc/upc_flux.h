// Flux-row layout and the per-node phi sum shared by the FP64 and MFMA
// luminosity forms.
#pragma once

#include "upc_internal.h"
#include "upc_qags.h"

#include <climits>

namespace upc {

struct FluxParams {
  double bmin, R, g1, g2, factor, twoR;
  int A, isPoint;
  int mirror_ny;  // > 0: symmetric grid with ny Y bins (k2 rows shared with mirror bins)
};

// Row holding k2 of point p.
__device__ __forceinline__ int k2_row(int p, int npts, int mirror_ny)
{
  if (mirror_ny <= 0) return npts + p;
  const int im = p / mirror_ny, iy = p - im * mirror_ny;
  return iy ? im * mirror_ny + (mirror_ny - iy) : npts + im;
}

__device__ __forceinline__ double row_node(double bmin, double logd, int j, double* bl, double* bh)
{
  *bl = bmin * exp(j * logd);
  *bh = bmin * exp((j + 1.) * logd);
  return (*bh + *bl) / 2.;
}

// Position guess on a row's log-spaced nodes b_j = (bmin e^{j logd} +
// bmin e^{(j+1) logd}) / 2 = c0 e^{j logd}: the first j with b_j >= v, from
// one float log2 (right up to +-1 where v lies within rounding of a node).
// Callers step from it to the exact index with their own comparisons.
struct LogGrid {
  float l2c0, s;  // log2(c0), ln 2 / logd
  __device__ __forceinline__ void init(double bmin, double logd)
  {
    l2c0 = (float)log2(bmin * (1.0 + exp(logd)) * 0.5);
    s = (float)(0.69314718055994531 / logd);
  }
  __device__ __forceinline__ int first_ge(double v, int n) const
  {
    if (!(v > 0.0)) return 0;
    const float g = ceilf((__log2f((float)v) - l2c0) * s);
    return g < 0.f ? 0 : (g > (float)n ? n : (int)g);
  }
};

__device__ __forceinline__ double node_b2(double b1, double b2, double c)
{
  return fma(2.0 * b1 * b2, c, fma(b1, b1, b2 * b2));
}

// The phi sum of one node pair: sum_k w_k P(min(b, 20)) G(b) (G = 1 for
// b >= 20); polarised: cos^2 and sin^2 weighted sums.
//
// Below the hadronic knee (b < ~11 fm for Pb-Pb) the G_AA spline holds only
// the decaying ringing of the natural spline, |G| < 1e-30.  Below bskip
// (x < bskip2) |G| < gskip = p20 2^-70 holds for every b, and P G is taken
// as 0 without sqrt, G or P: the caller subtracts p20 sum_k w_k from the
// same sum, and |P G| <= |G| is below 2^-70 of that term, 10^5 times below
// the rounding that the order of the pair sums already brings.  Above
// bskip, modes 2-4 still skip the breakup spline wherever the G just
// evaluated is below gskip (mode 1 has P = 1 and reads no breakup spline).
template <bool POL>
__device__ __forceinline__ void node_k(double b1, double b2, const double* cosk, const double* w5,
                                       int mode, double p20, double gskip, double bskip2, const double4* gaa,
                                       const double4* __restrict__ bk, double* ks, double* kp)
{
  const double db = 20. / (kNB - 1);
  double s = 0.0, p = 0.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double c = cosk[k];
    const double x = node_b2(b1, b2, POL ? -c : c);
    double breakup = mode != 1 ? p20 : 1., gaa_v = 1.;
    if (x < bskip2) {
      breakup = 0.0;
    } else if (x < 400.) {  // sqrt(x) >= 20 whenever x >= 400
      const double b = sqrt(x);
      if (b < 20.) {
        gaa_v = spline_eval_near(gaa, 0.0, db, 1.0 / db, kNB, b);
        if (mode != 1)
          breakup = fabs(gaa_v) < gskip ? 0.0 : spline_eval_near(bk, kBkMin, kDBk, 1.0 / kDBk, kNBkUsed, b);
      }
    }
    if (POL) {  // w5[5 + k] = sin^2(phi_k)
      s += breakup * gaa_v * w5[k] * c * c;
      p += breakup * gaa_v * w5[k] * w5[5 + k];
    } else {
      s += breakup * gaa_v * w5[k];
    }
  }
  *ks = s;
  *kp = p;
}


// fluxPoint (src/UpcCrossSection.cpp:166-178): g1 for both beams.
__device__ __forceinline__ double flux_point(double b, double k, double g1, double factor)
{
  const double g = g1;
  const double x = b * k / g / kHc;
  double K0 = 0, K1 = 0;
  if (x > 1e-10) bessel_k01(x, &K0, &K1);
  return factor * k / g / g * (K1 * K1 + K0 * K0 / g / g);
}

// Geometry of flux row r of a grid of npts points (src/UpcCrossSection.cpp:
// 224-233): photon energy, log step of the b grid, and the number of its
// nodes with b <= 2R (form-factor QAGS nodes; 0 for the point flux).
__device__ __forceinline__ void row_setup(const double* __restrict__ pts, int npts, int r, const FluxParams& fp,
                                          double* k_out, double* logd_out, int* cnt_out)
{
  int p, side;
  if (r < npts) {
    p = r;
    side = 0;
  } else {
    p = fp.mirror_ny > 0 ? (r - npts) * fp.mirror_ny : r - npts;
    side = 1;
  }
  const double M = pts[4 * p], Y = pts[4 * p + 1];
  const double k = side ? M / 2. * exp(-Y) : M / 2. * exp(Y);
  const double g = side ? fp.g2 : fp.g1;
  const double bmax = fmax(5. * g * kHc / k, 5. * fp.R);
  const double logd = (log(bmax) - log(fp.bmin)) / kNB1;
  int cnt = 0;
  if (!fp.isPoint) {
    // b_j increases with j: binary search for the number of nodes with b <= 2R
    int lo = 0, hi = kNB1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      double bl, bh;
      if (row_node(fp.bmin, logd, mid, &bl, &bh) > fp.twoR) hi = mid;
      else lo = mid + 1;
    }
    cnt = lo;
  }
  *k_out = k;
  *logd_out = logd;
  *cnt_out = cnt;
}

// fluxFormIntegrand (src/UpcCrossSection.cpp:181-191):
//   q^2 F(min(t, Q2max - dQ2)) / t * J1(b q / hc),  t = q^2 + (k / g1)^2
//
// The table's last knot is Q2min + (nQ2 - 1) dQ2 = 2 - 2e-6 (to rounding),
// so for t in (last knot, Q2max) the reference's gsl_spline_eval is out of
// range: GSL raises GSL_EDOM (its default handler aborts) and returns NaN.
// The integrand returns that NaN, so the QAGS run continues exactly as
// QUADPACK does with it, ends with a NaN result, and the integral is
// counted as failed, where the reference process would have stopped.
struct FluxFormIntegrand {
  const double4* __restrict__ ff;
  double b, w2g2;
  // The spline record is a dependent load from a 32 MB table: it is issued
  // first and J1 (independent ALU work) runs while it is in flight.  The
  // interval is the rounded estimate alone, as in spline_eval_near: within
  // rounding of a knot the neighbouring cubic agrees to a few ulp.
  __device__ __forceinline__ double operator()(double q) const
  {
    const double t = q * q + w2g2;
    const double tc = t < kQ2max ? t : kQ2max - kDQ2;
    int i = (int)((tc - kQ2min) * (1.0 / kDQ2));
    i = i < 0 ? 0 : (i > kNQ2 - 2 ? kNQ2 - 2 : i);
    const double4 c = ff[i];
    const double j1 = bessel_j1(b * q * (1.0 / kHc));
    const double d = tc - knot(kQ2min, kDQ2, i);
    const double f = c.x + d * (c.y + d * (c.z + d * c.w));
    const double v = q * q * f * rcp_nr(t) * j1;  // t >= (k / g1)^2 > 0
    return tc > kQ2Last ? __builtin_nan("") : v;
  }
};

// One form-factor QAGS task of a flux row: its node (b, k) and the flux slot
// (row * 120 + node) the integral fills.
struct TaskRec {
  double b, k;
  long long out;
};

// Node j of flux row r: the point-charge flux (b > 2R or FLUX_POINT 1), or a
// QAGS task record at the row's task offset.
__device__ __forceinline__ void point_row_node(long long r, int j, const FluxParams& fp,
                                               const double* __restrict__ row_k, const double* __restrict__ row_logd,
                                               const int* __restrict__ row_cnt,
                                               const long long* __restrict__ row_off, double* __restrict__ flux,
                                               TaskRec* __restrict__ recs)
{
  double bl, bh;
  const double b = row_node(fp.bmin, row_logd[r], j, &bl, &bh);
  if (j < row_cnt[r]) {
    TaskRec* t = recs + row_off[r] + j;
    t->b = b;
    t->k = row_k[r];
    t->out = r * kNB1 + j;
    return;
  }
  flux[r * kNB1 + j] = flux_point(b, row_k[r], fp.g1, fp.factor);
}

// One form-factor flux integral (fluxForm, src/UpcCrossSection.cpp:194-218):
// flux[out] = factor (I / A)^2 / k.
struct FluxTask {
  double b, k;
  long out;
  const double4* ff;  // form-factor spline of the task's nucleus
  double g1, factor;
  int A;
  int slot;  // job of a batched task (0 otherwise)
};

// Task sources of the QAGS kernel: get(t) describes task t.
struct GridTask {  // flux-row nodes of one context
  const TaskRec* __restrict__ recs;
  FluxParams fp;
  const double4* ff;
  __device__ void get(long long t, FluxTask* ft) const
  {
    const TaskRec rec = recs[t];
    ft->b = rec.b;
    ft->k = rec.k;
    ft->out = rec.out;
    ft->ff = ff;
    ft->g1 = fp.g1;
    ft->factor = fp.factor;
    ft->A = fp.A;
    ft->slot = 0;
  }
  // An integral that ended with ier != 0 (counted in the launch's stats).
  __device__ void failed(const FluxTask&) const {}
};

struct PointTask {  // arbitrary (b, k) points of one context
  const double* __restrict__ b;
  const double* __restrict__ k;
  const int* __restrict__ idx;  // compacted indices of points needing QAGS
  FluxParams fp;
  const double4* ff;
  __device__ void get(long long t, FluxTask* ft) const
  {
    const int i = idx[t];
    ft->b = b[i];
    ft->k = k[i];
    ft->out = i;
    ft->ff = ff;
    ft->g1 = fp.g1;
    ft->factor = fp.factor;
    ft->A = fp.A;
    ft->slot = 0;
  }
  __device__ void failed(const FluxTask&) const {}
};

// ---------------------------------------------------------------------------
// Form-factor flux: persistent QAGS kernel.  Each lane owns one integral at a
// time; lanes take new tasks from a global cursor as they finish, so the
// wavefront keeps issuing full-width QK21 passes.  stats: [0] integrand
// evaluations, [1] integrals with ier != 0, [2] overflow count.
// ---------------------------------------------------------------------------
// Waves per SIMD the QAGS kernel is compiled for (its register budget).
#ifndef UPC_QAGS_MIN_WAVES
#define UPC_QAGS_MIN_WAVES 4
#endif

// First QK21 pass of every task (the whole range [0, 10]): one task per lane
// at a time, every lane busy, results kept for k_flux_qags.
template <class TS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(UPC_QAGS_MIN_WAVES, 8))) void k_flux_first(
    TS ts, const unsigned long long* __restrict__ ntasks_p, double4* __restrict__ first)
{
  __shared__ double fvbuf[20 * 256];
  const long ntasks = (long)*ntasks_p;
  double* fv = fvbuf + (threadIdx.x >> 6) * 20 * 64 + (threadIdx.x & 63);
  const long stride = (long)gridDim.x * blockDim.x;
  FluxFormIntegrand f;
  for (long t0 = (long)blockIdx.x * blockDim.x; t0 < ntasks; t0 += stride) {
    const long t = t0 + (threadIdx.x & ~63) + (threadIdx.x & 63);
    const bool live = t < ntasks;
    FluxTask ft{};
    ts.get(live ? t : ntasks - 1, &ft);
    f.ff = ft.ff;
    f.b = ft.b;
    f.w2g2 = ft.k * ft.k / ft.g1 / ft.g1;
    const QK21Out o = qk21_lds(f, kQagsA, kQagsB, fv, 64);
    if (live) first[t] = make_double4(o.result, o.abserr, o.resabs, o.resasc);
  }
}

// first (may be null): the tasks' first QK21 passes from k_flux_first.  With
// it, a lane that takes a task applies that pass at once (finishing the
// integral when it already meets the tolerance), so every step in the loop
// below is a two-pass bisection step.
template <int CAP, class TS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(UPC_QAGS_MIN_WAVES, 8))) void k_flux_qags(
    TS ts, const unsigned long long* __restrict__ ntasks_p, unsigned long long* __restrict__ cursor,
    double* __restrict__ flux, int* __restrict__ ier_out, unsigned long long* __restrict__ stats,
    long long* __restrict__ overflow, int fast_cap, const double4* __restrict__ first)
{
  __shared__ double fvbuf[20 * 256];
  typename QagsSmallStore<CAP>::Arrays lists;
  Qags<QagsSmallStore<CAP>> q;
  q.s.a = &lists;
  q.s.lim = fast_cap < CAP ? fast_cap : CAP;
  const long ntasks = (long)*ntasks_p;
  const int lane = threadIdx.x & 63;
  double* fv = fvbuf + (threadIdx.x >> 6) * 20 * 64 + lane;
  long task = -1;  // -1: needs a task; LONG_MAX: exhausted
  FluxTask ft{};
  FluxFormIntegrand f;
  unsigned long long my_evals = 0, my_bad = 0;
  auto retire = [&](int st) {  // a finished or overflowing integral leaves the lane
    if (st == kQagsDone) {
      const double Q = q.result / ft.A;
      flux[ft.out] = ft.factor * Q * Q / ft.k;
      const int status = q.status();
      if (ier_out) ier_out[ft.out] = status;
      my_evals += (unsigned long long)q.neval();
      my_bad += status != 0;
      if (status != 0) ts.failed(ft);
      task = -1;
    } else if (st == kQagsOverflow) {
      const unsigned long long slot = atomicAdd(&stats[2], 1ull);
      overflow[slot] = task;
      task = -1;
    }
  };
  while (true) {
    bool need = task < 0;
    unsigned long long mask;
    while ((mask = __ballot(need)) != 0) {
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(cursor, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        const unsigned long long rank = __popcll(mask & ((1ull << lane) - 1ull));
        const long t = (long)(base + rank);
        if (t < ntasks) {
          task = t;
          ts.get(t, &ft);
          f.ff = ft.ff;
          f.b = ft.b;
          f.w2g2 = ft.k * ft.k / ft.g1 / ft.g1;
          q.start();
          need = false;
          if (first) {
            const double4 v = first[t];
            const QK21Out o{v.x, v.y, v.z, v.w};
            const int st = q.consume(o, o);
            retire(st);
            need = task < 0;
          }
        } else {
          task = LONG_MAX;
          need = false;
        }
      }
    }
    if (__ballot(task != LONG_MAX) == 0) break;
    if (task != LONG_MAX) {
      QK21Out o, o2{};
      const int np = q.passes();
#pragma unroll 1
      for (int k = 0; k < np; ++k) {  // one copy of the integrand for both passes
        double qa, qb;
        q.interval(k, &qa, &qb);
        const QK21Out r = qk21_lds(f, qa, qb, fv, 64);
        if (k == 0) o = r;
        else o2 = r;
      }
      retire(q.consume(o, o2));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    my_evals += __shfl_down(my_evals, o);
    my_bad += __shfl_down(my_bad, o);
  }
  if (lane == 0) {
    atomicAdd(&stats[0], my_evals);
    atomicAdd(&stats[1], my_bad);
  }
}

// Tasks that outgrew the fast interval list (integrals that bisect more than
// the fast list holds: in practice the ones whose integrand turned NaN or whose
// roundoff tests fire, which run on to dozens or hundreds of intervals) are
// rerun from the start here, one wavefront per integral, with the reference's
// 1000-interval workspace in LDS.  Lanes 0..41 evaluate the 42 nodes of a
// step (both halves) at once; lane 0 sums each pass in QK21's order (the same
// qk21_sum as the fast kernel) and runs the dqagse state machine, so the
// serial part of a step is LDS work and one sum per pass.
constexpr int kQagsWaveBlocks = 1024;  // workgroups of the overflow kernel (36.5 KB LDS each)

template <class TS>
__global__ __launch_bounds__(64) void k_flux_qags_wave(TS ts, const long long* __restrict__ overflow,
                                                       unsigned long long* __restrict__ stats,
                                                       double* __restrict__ flux, int* __restrict__ ier_out)
{
  __shared__ QagsBigArrays ws;
  __shared__ double vals[2][21];
  __shared__ double fvs[20];
  __shared__ double iv[2][2];
  __shared__ int ctl[2];  // [0] passes of the step, [1] state after it
  const long n = (long)stats[2];
  const int lane = threadIdx.x;
  for (long i = blockIdx.x; i < n; i += gridDim.x) {
    FluxTask ft;
    ts.get(overflow[i], &ft);
    FluxFormIntegrand f;
    f.ff = ft.ff;
    f.b = ft.b;
    f.w2g2 = ft.k * ft.k / ft.g1 / ft.g1;
    Qags<QagsBigStore> q;
    q.s.a = &ws;
    if (lane == 0) q.start();
    int st = kQagsRunning;
    while (true) {
      if (lane == 0) {
        const int np = q.passes();
        ctl[0] = np;
        for (int k = 0; k < np; ++k) q.interval(k, &iv[k][0], &iv[k][1]);
      }
      __syncthreads();
      const int np = ctl[0];
      if (lane < 21 * np) {
        const int pass = lane >= 21, e = lane - 21 * pass;
        const double a = iv[pass][0], b = iv[pass][1];
        vals[pass][e] = f(qk21_abscissa(e, 0.5 * (a + b), 0.5 * (b - a)));
      }
      __syncthreads();
      if (lane == 0) {
        const QK21Out o = qk21_sum([&](int e, double) { return vals[0][e]; }, iv[0][0], iv[0][1], fvs, 1);
        QK21Out o2{};
        if (np == 2) o2 = qk21_sum([&](int e, double) { return vals[1][e]; }, iv[1][0], iv[1][1], fvs, 1);
        ctl[1] = q.consume(o, o2);
      }
      __syncthreads();
      st = ctl[1];
      if (st != kQagsRunning) break;
    }
    if (lane == 0) {
      const double Q = q.result / ft.A;
      flux[ft.out] = ft.factor * Q * Q / ft.k;
      const int status = st == kQagsDone ? q.status() : 99;
      if (ier_out) ier_out[ft.out] = status;
      atomicAdd(&stats[0], (unsigned long long)q.neval());
      if (status != 0) {
        atomicAdd(&stats[1], 1ull);
        ts.failed(ft);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// FP64 luminosity: one wavefront per (M, Y) point.
//
// For a pair of nodes (b1_i, b2_j) the phi sum K_ij = sum_k w_k P(b) G(b),
// b^2 = b1^2 + b2^2 +- 2 b1 b2 cos(phi_k).  Where every node has b >= 20 fm
// the reference evaluates P at the 20 fm clamp and G = 1, so K_ij is the
// constant S0 (same arithmetic).  For fixed i the pairs with some b < 20 form
// one contiguous j range (the smallest b^2 is always the extreme-cos node), so
//   L = 2 pi^2 M [ S0 (sum_i W1_i)(sum_j W2_j) + sum_{band} W1_i W2_j (K_ij - S0) ]
// with W = N(b, k) b (b_h - b_l).  The band pairs are spread evenly over the
// 64 lanes.
// ---------------------------------------------------------------------------

// What the contraction reads of one job (one context's grid).
struct LumiView {
  const double* pts;       // (M, Y, dm, dy) of the job's points
  const double* row_logd;  // the job's flux rows (k2 rows after the k1 rows)
  const double* flux;
  FluxParams fp;
  const double4* gaa;
  const double4* bk;
  const DevScalars* sc;
  int npts, mode, pol;
  double* L;
  double* Ls;
  double* Lp;
  // 0: one work item per point.  1 (symmetric grid, fp.mirror_ny = ny > 0):
  // bins (m, iy) and (m, ny - iy) have the same M, dm, dy and exchanged flux
  // rows, and the node sum is symmetric under b1 <-> b2, so one evaluation
  // fills both; work items are the bins iy <= ny / 2.
  int mirror_out;
};

// Work items of a contraction launch.
__host__ __device__ inline int lumi_work_items(int npts, int mirror_ny, int mirror_out)
{
  return mirror_out ? npts / mirror_ny * (mirror_ny / 2 + 1) : npts;
}

// Work item q -> point p and its mirror bin pm (-1: none).
__device__ __forceinline__ void work_point(int q, int mirror_ny, int mirror_out, int* p, int* pm)
{
  if (!mirror_out) {
    *p = q;
    *pm = -1;
    return;
  }
  const int ny = mirror_ny, nc = ny / 2 + 1;
  const int im = q / nc, iy = q - im * nc;
  *p = im * ny + iy;
  *pm = iy > 0 && ny - iy != iy ? im * ny + (ny - iy) : -1;
}

__device__ __forceinline__ void lumi_work_point(const LumiView& v, int q, int* p, int* pm)
{
  work_point(q, v.fp.mirror_ny, v.mirror_out, p, pm);
}

// Job resolvers: the job and first point of a block.
struct SingleLumi {
  LumiView v;
  __device__ void get(int blk, int waves, LumiView* out, int* p0) const
  {
    *out = v;
    *p0 = blk * waves;
  }
};

template <int WAVES>
struct LumiSmem {
  double4 gaa[kNB];
  double sb1[WAVES][kNB1], sb2[WAVES][kNB2];
  double sw1[WAVES][kNB1], sw2[WAVES][kNB2];
  int spre[WAVES][kNB1 + 1], sjlo[WAVES][kNB1];
  double sp2[WAVES][kNB2 + 1];  // prefix sums of sw2
  double cosk[5], w5[10];
};

template <bool POL, int WAVES>
__device__ __forceinline__ void lumi_fp64_body(const LumiView& v, int p0, LumiSmem<WAVES>& sm)
{
  for (int i = threadIdx.x; i < kNB - 1; i += blockDim.x) sm.gaa[i] = v.gaa[i];
  if (threadIdx.x < 5) {
    const double c = cos(kPi * kGLx5[threadIdx.x]);
    const double sn = sin(kPi * kGLx5[threadIdx.x]);
    sm.cosk[threadIdx.x] = c;
    sm.w5[threadIdx.x] = kGLw5[threadIdx.x];
    sm.w5[5 + threadIdx.x] = sn * sn;
  }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = p0 + wv;
  const bool live = q < lumi_work_items(v.npts, v.fp.mirror_ny, v.mirror_out);
  int p = 0, pm = -1;
  if (live) lumi_work_point(v, q, &p, &pm);
  const FluxParams& fp = v.fp;
  const double p20 = v.sc->p20;
  double M = 0, dm = 1, dy = 1;
  LogGrid lg2{};
  if (live) {
    M = v.pts[4 * p];
    dm = v.pts[4 * p + 2];
    dy = v.pts[4 * p + 3];
    const int r2 = k2_row(p, v.npts, fp.mirror_ny);
    const double logd1 = v.row_logd[p], logd2 = v.row_logd[r2];
    lg2.init(fp.bmin, logd2);
    const double* f1 = v.flux + (long)p * kNB1;
    const double* f2 = v.flux + (long)r2 * kNB1;
    for (int j = lane; j < kNB1; j += 64) {
      double bl, bh;
      const double b1 = row_node(fp.bmin, logd1, j, &bl, &bh);
      sm.sb1[wv][j] = b1;
      sm.sw1[wv][j] = f1[j] * b1 * (bh - bl);
      const double b2 = row_node(fp.bmin, logd2, j, &bl, &bh);
      sm.sb2[wv][j] = b2;
      sm.sw2[wv][j] = f2[j] * b2 * (bh - bl);
    }
  }
  __syncthreads();
  // the extreme-cos node has the smallest b for every pair, cmax the largest
  const double cext = POL ? -sm.cosk[0] : sm.cosk[4];
  const double cmax = POL ? -sm.cosk[4] : sm.cosk[0];
  if (!live) return;
  // prefix sums of sw2: sp2[j] = sum_{j' < j} sw2[j']
  {
    double a0 = sm.sw2[wv][lane], a1 = lane + 64 < kNB2 ? sm.sw2[wv][lane + 64] : 0.0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double u0 = __shfl_up(a0, o), u1 = __shfl_up(a1, o);
      if (lane >= o) {
        a0 += u0;
        a1 += u1;
      }
    }
    const double half = __shfl(a0, 63);
    sm.sp2[wv][lane + 1] = a0;
    if (lane + 64 < kNB2) sm.sp2[wv][lane + 65] = half + a1;
    if (lane == 0) sm.sp2[wv][0] = 0.0;
  }
  const double bskip = v.sc->bskip, bskip2 = bskip * bskip;
  // Pairs whose largest node is below bskip have P G = 0 at every node, so
  // they add -s0 sw1[i] sw2[j]: per row, -s0 sw1[i] sp2[ja] for the run
  // j < ja (b^2 at cmax > 0 grows with b2).  Those pairs are in the band
  // (b^2 < bskip^2 <= 400 at every node), which therefore starts at j = 0;
  // the walk takes the band from ja on.
  double wan = 0.0;  // sum over rows of sw1[i] sp2[ja]
  wave_sync_lds();
  // Band of row i: the j with b^2 < 400 at the extreme cosine.  b^2 is convex
  // in b2 (smallest at b2 = -b1 cext), so the band is the run of j around
  // that vertex.  The vertex and both ends (roots b2 = vtx -+ sqrt(400 -
  // b1^2 (1 - cext^2))) are guessed on the log grid of b2 and stepped to
  // exactness with the same comparisons a bisection would make.
  int cnt[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = lane + 64 * h;
    cnt[h] = 0;
    if (i < kNB1) {
      const double b1 = sm.sb1[wv][i];
      const double* b2s = sm.sb2[wv];
      auto inb = [&](int j) { return node_b2(b1, b2s[j], cext) < 400.0; };
      const double vtx = -b1 * cext;
      int lo = lg2.first_ge(vtx, kNB2);  // first j with b2_j >= vtx
      while (lo > 0 && b2s[lo - 1] >= vtx) --lo;
      while (lo < kNB2 && b2s[lo] < vtx) ++lo;
      int jc = -1;
      if (lo < kNB2 && inb(lo)) jc = lo;
      if (lo > 0 && inb(lo - 1)) jc = lo - 1;
      int jlo = 0;
      if (jc >= 0) {
        const double d = 400.0 - b1 * b1 * (1.0 - cext * cext), sd = sqrt(d > 0.0 ? d : 0.0);
        int a = lg2.first_ge(vtx - sd, kNB2);  // first j in band: outside before it, inside from it to jc
        a = a > jc ? jc : a;
        while (a > 0 && inb(a - 1)) --a;
        while (a < jc && !inb(a)) ++a;
        jlo = a;
        a = lg2.first_ge(vtx + sd, kNB2) - 1;  // last j in band
        a = a < jc ? jc : (a > kNB2 - 1 ? kNB2 - 1 : a);
        while (a < kNB2 - 1 && inb(a + 1)) ++a;
        while (a > jc && !inb(a)) --a;
        // first j with b^2 >= bskip^2 at cmax (b^2 grows with b2 there)
        auto low = [&](int j) { return node_b2(b1, b2s[j], cmax) < bskip2; };
        const double da = bskip2 - b1 * b1 * (1.0 - cmax * cmax);
        int ja = da > 0.0 ? lg2.first_ge(-b1 * cmax + sqrt(da), kNB2) : 0;
        while (ja > 0 && !low(ja - 1)) --ja;
        while (ja < kNB2 && low(ja)) ++ja;
        if (ja > 0) {
          wan += sm.sw1[wv][i] * sm.sp2[wv][ja];
          jlo = ja;
        }
        cnt[h] = a >= jlo ? a - jlo + 1 : 0;
      }
      sm.sjlo[wv][i] = jlo;
    }
  }
  // inclusive wave scan of the two halves of the row counts
  int c0 = cnt[0], c1 = cnt[1];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u0 = __shfl_up(c0, o), u1 = __shfl_up(c1, o);
    if (lane >= o) {
      c0 += u0;
      c1 += u1;
    }
  }
  const int half = __shfl(c0, 63);
  sm.spre[wv][lane + 1] = c0;
  if (lane + 64 < kNB1) sm.spre[wv][lane + 65] = half + c1;
  if (lane == 0) sm.spre[wv][0] = 0;
  wave_sync_lds();
  // S0: the phi sum when every node is at or beyond 20 fm
  double s0s = 0, s0p = 0;
  {
    const double br = v.mode != 1 ? p20 : 1.;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (POL) {
        s0s += br * 1. * sm.w5[k] * sm.cosk[k] * sm.cosk[k];
        s0p += br * 1. * sm.w5[k] * sm.w5[5 + k];
      } else {
        s0s += br * 1. * sm.w5[k];
      }
    }
  }
  const int T = sm.spre[wv][kNB1];
  const double gskip = p20 * 0x1p-70;
  double accs = 0, accp = 0;
  int i = 0;  // row of pair t: pre[i] <= t < pre[i + 1]; t only grows
  for (int t = lane; t < T; t += 64) {
    while (sm.spre[wv][i + 1] <= t) ++i;
    const int j = sm.sjlo[wv][i] + (t - sm.spre[wv][i]);
    double ks, kp;
    node_k<POL>(sm.sb1[wv][i], sm.sb2[wv][j], sm.cosk, sm.w5, v.mode, p20, gskip, bskip2, sm.gaa, v.bk, &ks,
                &kp);
    const double ww = sm.sw1[wv][i] * sm.sw2[wv][j];
    accs += ww * (ks - s0s);
    if (POL) accp += ww * (kp - s0p);
  }
  double sw1t = 0, sw2t = 0;
  for (int j = lane; j < kNB1; j += 64) {
    sw1t += sm.sw1[wv][j];
    sw2t += sm.sw2[wv][j];
  }
  for (int o = 32; o > 0; o >>= 1) {
    accs += __shfl_down(accs, o);
    accp += __shfl_down(accp, o);
    sw1t += __shfl_down(sw1t, o);
    sw2t += __shfl_down(sw2t, o);
    wan += __shfl_down(wan, o);
  }
  if (lane == 0) {
    const double pre = 2 * kPi * kPi * M;
    const double tot_s = s0s * (sw1t * sw2t - wan) + accs;
    if (POL) {
      const double tot_p = s0p * (sw1t * sw2t - wan) + accp;
      const double ls = pre * tot_s * dm * dy, lp = pre * tot_p * dm * dy;
      v.Ls[p] = ls;
      v.Lp[p] = lp;
      if (pm >= 0) {
        v.Ls[pm] = ls;
        v.Lp[pm] = lp;
      }
    } else {
      const double l = pre * tot_s * dm * dy;
      v.L[p] = l;
      if (pm >= 0) v.L[pm] = l;
    }
  }
}

// WAVES points per block; the resolver maps the block to its job (all
// threads of a block take the same branch, so the barriers are uniform).
template <int WAVES, class JR>
__global__ __launch_bounds__(64 * WAVES) void k_lumi_fp64(JR jr)
{
  __shared__ LumiSmem<WAVES> sm;
  LumiView v;
  int p0;
  jr.get(blockIdx.x, WAVES, &v, &p0);
  if (v.pol) lumi_fp64_body<true, WAVES>(v, p0, sm);
  else lumi_fp64_body<false, WAVES>(v, p0, sm);
}

// Three-pass exclusive scan of the row counts: per-block sums, their scan in
// one block, then the offsets.  *total receives the sum.
constexpr int kScanBlock = 1024;

template <int BS = kScanBlock>
__global__ __launch_bounds__(BS) void k_scan_part(const int* __restrict__ cnt, long n,
                                                          long long* __restrict__ part)
{
  __shared__ long long s[BS];
  const long i = (long)blockIdx.x * BS + threadIdx.x;
  s[threadIdx.x] = i < n ? cnt[i] : 0;
  __syncthreads();
  for (int d = BS / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

template <int BS = kScanBlock>
__global__ __launch_bounds__(BS) void k_scan_top(long long* __restrict__ part, int nparts,
                                                         unsigned long long* __restrict__ total)
{
  __shared__ long long s[BS];
  long long carry = 0;
  for (int base = 0; base < nparts; base += BS) {
    const int i = base + threadIdx.x;
    const long long v = i < nparts ? part[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < BS; d <<= 1) {
      const long long t = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nparts) part[i] = carry + s[threadIdx.x] - v;  // exclusive
    const long long blk_total = s[BS - 1];
    __syncthreads();
    carry += blk_total;
  }
  if (threadIdx.x == 0) *total = (unsigned long long)carry;
}

template <int BS = kScanBlock>
__global__ __launch_bounds__(BS) void k_scan_apply(const int* __restrict__ cnt, long n,
                                                           const long long* __restrict__ part,
                                                           long long* __restrict__ off)
{
  __shared__ long long s[BS];
  const long i = (long)blockIdx.x * BS + threadIdx.x;
  const long long v = i < n ? cnt[i] : 0;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int d = 1; d < BS; d <<= 1) {
    const long long t = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) off[i] = part[blockIdx.x] + s[threadIdx.x] - v;
}

inline FluxParams flux_params(const upc_ctx* c, int mirror_ny = 0)
{
  FluxParams fp;
  fp.mirror_ny = mirror_ny;
  fp.R = c->cfg.R;
  fp.bmin = c->cfg.flux_point ? c->cfg.R : 0.05 * c->cfg.R;
  fp.g1 = c->g1;
  fp.g2 = c->g2;
  fp.factor = c->factor;
  fp.twoR = 2. * c->cfg.R;
  fp.A = c->cfg.A;
  fp.isPoint = c->cfg.flux_point;
  return fp;
}

// Builds the flux rows of the points in the work buffer (d_pts == nullptr)
// or of the npts (M, Y) pairs at d_pts ([M..., Y...]): point-charge nodes
// directly, form-factor nodes by the QAGS task kernels.  *fp receives the
// row parameters (mirror_ny set when the grid's Y range is symmetric).
int prepare_flux(upc_ctx* c, const double* d_pts, int npts, FluxParams* fp, hipStream_t s);
// Copies the QAGS counters of the last prepare_flux into the context.
int read_flux_stats(upc_ctx* c, hipStream_t s);

}  // namespace upc
