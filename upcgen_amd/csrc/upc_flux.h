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
// range: GSL raises GSL_EDOM (its default handler ends the program).  The
// integrand returns NaN there and the QAGS run stops after that QK21 pass
// (kQagsEdom).
//
// The spline record is a dependent load from a 32 MB table: it is issued
// first and J1 (independent ALU work) runs while it is in flight.  The
// interval is the rounded estimate alone, as in spline_eval_near: within
// rounding of a knot the neighbouring cubic agrees to a few ulp.
//
// BR fixes J1's branch: 0 lets J1 choose (x <= 8 or not), 1 and 2 take the
// x <= 8 and x > 8 branches for a caller that has sorted its nodes by
// flux_j1_arg.  The operations are the same in all three forms, so a value
// does not depend on which one computed it.
__device__ __forceinline__ double flux_j1_arg(double b, double q) { return b * q * (1.0 / kHc); }

// The integrand in two steps: the spline record's load (issued early) and
// the rest once J1 is known.
struct FluxNode {
  double q, t, tc;
  int i;
  double4 c;  // spline interval record (in flight after flux_prelude)
};

__device__ __forceinline__ void flux_prelude(const double4* __restrict__ ff, double w2g2, double q, FluxNode* n)
{
  n->q = q;
  n->t = q * q + w2g2;
  n->tc = n->t < kQ2max ? n->t : kQ2max - kDQ2;
  int i = (int)((n->tc - kQ2min) * (1.0 / kDQ2));
  n->i = i < 0 ? 0 : (i > kNQ2 - 2 ? kNQ2 - 2 : i);
  n->c = ff[n->i];
}

__device__ __forceinline__ double flux_finish(const FluxNode& n, double j1)
{
  const double d = n.tc - knot(kQ2min, kDQ2, n.i);
  const double f = n.c.x + d * (n.c.y + d * (n.c.z + d * n.c.w));
  const double v = n.q * n.q * f * rcp_nr(n.t) * j1;  // t >= (k / g1)^2 > 0
  return n.tc > kQ2Last ? __builtin_nan("") : v;
}

template <int BR>
__device__ __forceinline__ double bessel_j1_branch(double x)
{
  return BR == 1 ? bessel_j1_small(x) : (BR == 2 ? bessel_j1_large(x) : bessel_j1(x));
}

template <int BR>
__device__ __forceinline__ double flux_integrand(const double4* __restrict__ ff, double b, double w2g2, double q)
{
  FluxNode n;
  flux_prelude(ff, w2g2, q, &n);
  return flux_finish(n, bessel_j1_branch<BR>(flux_j1_arg(b, q)));
}

struct FluxFormIntegrand {
  const double4* __restrict__ ff;
  double b, w2g2;
  __device__ __forceinline__ double operator()(double q) const { return flux_integrand<0>(ff, b, w2g2, q); }
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

// What the end of an integral needs: flux[out] = factor (I / A)^2 / k.
struct FluxOut {
  double k, factor;
  long out;
  int A;
};

// Task sources of the QAGS kernel: get(t) describes task t; out(t, job)
// gives its output fields again (job: FluxTask::slot of the same task).
struct GridTask {  // flux-row nodes of one context
  static constexpr bool kSharedFF = true;  // one form-factor table for every task (ff)
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
  __device__ void out(long long t, int, FluxOut* fo) const
  {
    const TaskRec rec = recs[t];
    fo->k = rec.k;
    fo->out = rec.out;
    fo->factor = fp.factor;
    fo->A = fp.A;
  }
  // An integral of job `job` that ended with ier != 0 (counted in the launch's stats).
  __device__ void failed(int) const {}
  __device__ const double4* any_ff() const { return ff; }
};

struct PointTask {  // arbitrary (b, k) points of one context
  static constexpr bool kSharedFF = true;
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
  __device__ void out(long long t, int, FluxOut* fo) const
  {
    const int i = idx[t];
    fo->k = k[i];
    fo->out = i;
    fo->factor = fp.factor;
    fo->A = fp.A;
  }
  __device__ void failed(int) const {}
  __device__ const double4* any_ff() const { return ff; }
};

// ---------------------------------------------------------------------------
// Form-factor flux: persistent QAGS kernel.
//
// Each lane owns one integral at a time (the dqagse state machine of
// upc_qags.h, its interval lists in private memory) and takes a new task
// from a global cursor when it finishes.  A loop iteration is one dqagse
// step for every lane: the first QK21 pass of a new task, or the two halves
// of a bisection, in at most two rounds of one pass per lane.
//
// A round evaluates the 21 nodes of every lane's pass as one pool: each lane
// writes its abscissae to its row of the wavefront's LDS table and queues
// its nodes as J1 small-argument items (x = b q / hc <= 8, front of the
// queue) or large-argument items (after them).  The wavefront then sweeps
// the queue 64 items at a time, J1's small branch over the front and its
// large branch over the rest, each lane evaluating whichever node the queue
// hands it (the owner's b and (k/g)^2 come over ds_bpermute) and writing
// the value over the abscissa.  Every sweep runs one J1 branch on all 64
// lanes, where evaluating each lane's own nodes in turn would run both
// branches on almost every node (the two kinds are about evenly mixed in a
// grid: 47% small, 53% large on the 1001 x 121 Pb-Pb grid).  Finally each
// lane sums its own row in QK21's order (qk21_stored), so results are the
// same bits as a lane evaluating its own nodes.
//
// The LDS table (13,440 B per wavefront) sets the occupancy: 3 workgroups
// of 4 wavefronts per CU, 168 VGPRs per lane.
//
// stats: [0] integrand evaluations, [1] integrals with ier != 0, [2]
// overflow count.
// ---------------------------------------------------------------------------
struct QagsWaveLds {
  double v[64 * 21];             // [owner lane][node e]: abscissa, then the integrand value
  unsigned short item[21 * 64];  // owner lane << 5 | e: small-argument items first, large after
};
constexpr int kQagsBlock = 256;      // threads of a workgroup (4 wavefronts, 53,760 B of LDS)
constexpr int kQagsBlocksPerCu = 3;  // resident workgroups per CU (LDS-bound)

__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

__device__ __forceinline__ int lanes_below(unsigned long long m)
{
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Items [beg, end) of the queue with J1 branch BR, 64 per sweep.  beg, end
// are wavefront-uniform, so every lane runs every ds_bpermute; a lane past
// end takes a dummy item (its own first node, abscissa 0) and stores
// nothing.  The fetch has no branch: every lane reads the queue (at a
// clamped index), its abscissa and a spline record (of a valid table: a
// lane's table pointer is set before its first task), so the record loads
// of a sweep's two items are issued unconditionally and each J1 waits only
// for its own record.
struct SweepItem {
  bool act;
  int slot;      // owner * 21 + e
  double x;      // J1 argument
  FluxNode n;
};

template <bool SHARED_FF>
__device__ __forceinline__ void sweep_fetch(QagsWaveLds& W, int lane, int r0, int end, const double4* ff, double b,
                                            double w2g2, SweepItem* it)
{
  constexpr int kItems = 21 * 64;
  const int r = r0 + lane;
  it->act = r < end;
  const int raw = W.item[r < kItems ? r : kItems - 1];
  const int code = it->act ? raw : lane << 5;
  const int owner = code >> 5, e = code & 31;
  const double bo = __shfl(b, owner), wo = __shfl(w2g2, owner);
  const double4* ffo = ff;
  if (!SHARED_FF) ffo = (const double4*)__shfl((long long)ff, owner);
  it->slot = owner * 21 + e;
  const double qv = W.v[it->slot];
  const double q = it->act ? qv : 0.0;
  it->x = flux_j1_arg(bo, q);
  flux_prelude(ffo, wo, q, &it->n);
}

template <int BR>
__device__ __forceinline__ void sweep_finish(QagsWaveLds& W, const SweepItem& it)
{
  const double v = flux_finish(it.n, bessel_j1_branch<BR>(it.x));
  if (it.act) W.v[it.slot] = v;
}

// Two items per lane in flight: both fetches (queue entry, owner's
// parameters, abscissa, spline record load) are issued before either J1.
// The J1s run at a raised wave priority, so the SIMD issues them ahead of
// the other waves' fetches and bookkeeping, whose waits then overlap them
// (2 % on the QAGS phase).  A remainder of at most 64 items takes one sweep
// of its own.
template <int BR, bool SHARED_FF>
__device__ __forceinline__ void qags_sweeps(QagsWaveLds& W, int lane, int beg, int end, const double4* ff,
                                            double b, double w2g2)
{
  int r0 = beg;
#pragma unroll 1
  for (; r0 + 64 < end; r0 += 128) {
    SweepItem ia, ib;
    sweep_fetch<SHARED_FF>(W, lane, r0, end, ff, b, w2g2, &ia);
    sweep_fetch<SHARED_FF>(W, lane, r0 + 64, end, ff, b, w2g2, &ib);
    __builtin_amdgcn_s_setprio(2);
    sweep_finish<BR>(W, ia);
    sweep_finish<BR>(W, ib);
    __builtin_amdgcn_s_setprio(0);
  }
  if (r0 < end) {
    SweepItem ia;
    sweep_fetch<SHARED_FF>(W, lane, r0, end, ff, b, w2g2, &ia);
    sweep_finish<BR>(W, ia);
  }
}

// Node of QK21 in ascending abscissa order k = 0..20: the left nodes
// centr - hlgth kXgk[k] (kXgk decreasing), the centre, the right nodes.
__device__ __forceinline__ int qk21_qorder_node(int k)
{
  return k < 10 ? qk21_node(k, 0) : (k == 10 ? 0 : qk21_node(20 - k, 1));
}

// Exclusive prefix sum over the wavefront's lanes of x in [0, 32), and the
// total: one ballot per bit.
__device__ __forceinline__ int wave_prefix5(int x, int* total)
{
  int pre = 0, tot = 0;
#pragma unroll
  for (int bit = 0; bit < 5; ++bit) {
    const unsigned long long m = __ballot((x >> bit) & 1);
    pre += lanes_below(m) << bit;
    tot += __popcll(m) << bit;
  }
  *total = tot;
  return pre;
}

// One QK21 pass per lane with mine set, on [qa, qb]; the result for those
// lanes (zeros elsewhere).  ff, b, w2g2: the lane's own integrand.
//
// J1's argument b q / hc grows with the abscissa, so a lane's small-argument
// nodes are the first ns of its nodes in abscissa order: the lane's items
// are two runs, [ps, ps + ns) among the small items and the rest among the
// large ones, placed by one prefix sum over the lanes.
template <bool SHARED_FF>
__device__ __forceinline__ QK21Out qk21_wave(QagsWaveLds& W, int lane, bool mine, double qa, double qb,
                                             const double4* ff, double b, double w2g2)
{
  const double centr = 0.5 * (qa + qb), hlgth = 0.5 * (qb - qa);
  double* col = W.v + lane * 21;
  wave_lds_sync();
  // the abscissae of qk21_abscissa, a pair (both sides of kXgk[j]) at a time
  col[0] = centr;
  int ns = flux_j1_arg(b, centr) <= 8.0;
#pragma unroll 5
  for (int j = 0; j < 10; ++j) {
#pragma clang fp contract(off)  // absc rounded before the sums, as QUADPACK computes it
    const double absc = hlgth * kXgk[j];
    const double ql = centr - absc, qr = centr + absc;
    const int e = qk21_node(j, 0);
    col[e] = ql;
    col[e + 1] = qr;
    ns += (flux_j1_arg(b, ql) <= 8.0) + (flux_j1_arg(b, qr) <= 8.0);
  }
  if (!mine) ns = 0;
  int S = 0;
  const int ps = wave_prefix5(ns, &S);
  const unsigned long long act = __ballot(mine);
  const int ntot = 21 * __popcll(act);
  const int pl = S + 21 * lanes_below(act) - ps;  // start of the lane's large run
  if (mine) {
#pragma unroll 1
    for (int k = 0; k < 21; ++k) {
      const int e = qk21_qorder_node(k);
      W.item[k < ns ? ps + k : pl + (k - ns)] = (unsigned short)(lane << 5 | e);
    }
  }
  wave_lds_sync();
  qags_sweeps<1, SHARED_FF>(W, lane, 0, S, ff, b, w2g2);
  qags_sweeps<2, SHARED_FF>(W, lane, S, ntot, ff, b, w2g2);
  wave_lds_sync();
  QK21Out o{};
  if (mine) o = qk21_stored(col, 1, qa, qb);
  return o;
}

template <int CAP, class TS>
__global__ __launch_bounds__(kQagsBlock) __attribute__((amdgpu_waves_per_eu(kQagsBlocksPerCu, kQagsBlocksPerCu)))
void k_flux_qags(TS ts, const unsigned long long* __restrict__ ntasks_p, unsigned long long* __restrict__ cursor,
                 double* __restrict__ flux, int* __restrict__ ier_out, unsigned long long* __restrict__ stats,
                 long long* __restrict__ overflow, int fast_cap)
{
  __shared__ QagsWaveLds lds[kQagsBlock / 64];
  QagsWaveLds& W = lds[threadIdx.x >> 6];
  typename QagsSmallStore<CAP>::Arrays lists;
  Qags<QagsSmallStore<CAP>> q;
  q.s.a = &lists;
  q.s.lim = fast_cap < CAP ? fast_cap : CAP;
  const long long ntasks = (long long)*ntasks_p;
  const int lane = threadIdx.x & 63;
  long long task = -1;  // -1: needs a task; LLONG_MAX: exhausted
  // Live across the steps of a task: b, (k/g)^2, the form-factor table and the
  // job; the output fields (k, slot, factor, A) are read again when it ends.
  const double4* ft_ff = ts.any_ff();  // a valid table for the sweeps' idle lanes before the first task
  double ft_b = 0.0, w2g2 = 0.0;
  int ft_job = 0;
  unsigned my_evals = 0, my_bad = 0;  // 32-bit per lane; evaluations flushed before they could wrap
  while (true) {
    bool need = task < 0;
    unsigned long long mask;
    while ((mask = __ballot(need)) != 0) {
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(cursor, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        const long long t = (long long)(base + (unsigned long long)lanes_below(mask));
        if (t < ntasks) {
          task = t;
          FluxTask ft;
          ts.get(t, &ft);
          ft_b = ft.b;
          ft_ff = ft.ff;
          ft_job = ft.slot;
          w2g2 = ft.k * ft.k / ft.g1 / ft.g1;
          q.start();
        } else {
          task = LLONG_MAX;
        }
        need = false;
      }
    }
    if (__ballot(task != LLONG_MAX) == 0) break;
    const bool live = task != LLONG_MAX;
    const int np = live ? q.passes() : 0;
    const double4* ff;
    if constexpr (TS::kSharedFF) ff = ts.ff;
    else ff = ft_ff;
    double qa = 0.0, qb = 0.0;
    if (live) q.interval(0, &qa, &qb);
    const QK21Out o = qk21_wave<TS::kSharedFF>(W, lane, live, qa, qb, ff, ft_b, w2g2);
    QK21Out o2{};
    if (__ballot(np == 2) != 0) {
      if (np == 2) q.interval(1, &qa, &qb);
      o2 = qk21_wave<TS::kSharedFF>(W, lane, np == 2, qa, qb, ff, ft_b, w2g2);
    }
    if (!live) continue;
    const int st = q.consume(o, o2);
    if (st == kQagsDone) {
      FluxOut fo;
      ts.out(task, ft_job, &fo);
      const double Q = q.result / fo.A;
      flux[fo.out] = fo.factor * Q * Q / fo.k;
      const int status = q.status();
      if (ier_out) ier_out[fo.out] = status;
      my_evals += (unsigned)q.neval();
      if (my_evals >= 0x80000000u) {
        atomicAdd(&stats[0], (unsigned long long)my_evals);
        my_evals = 0;
      }
      my_bad += status != 0;
      if (status != 0) ts.failed(ft_job);
      task = -1;
    } else if (st == kQagsOverflow) {
      const unsigned long long slot = atomicAdd(&stats[2], 1ull);
      overflow[slot] = task;
      task = -1;
    }
  }
  unsigned long long evals = my_evals, bad = my_bad;
  for (int o = 32; o > 0; o >>= 1) {
    evals += __shfl_down(evals, o);
    bad += __shfl_down(bad, o);
  }
  if (lane == 0) {
    atomicAdd(&stats[0], evals);
    atomicAdd(&stats[1], bad);
  }
}

// Tasks that outgrew the fast interval list (integrals that bisect more than
// the fast list holds: in practice the ones whose integrand turned NaN or whose
// roundoff tests fire, which run on to dozens or hundreds of intervals) are
// rerun from the start here, one wavefront per integral, with the reference's
// 1000-interval workspace in LDS.  Lanes 0..41 evaluate the 42 nodes of a
// step (both halves) at once; lane 0 sums each pass in QK21's order (the same
// qk21_core as the fast kernel) and runs the dqagse state machine, so the
// serial part of a step is LDS work and one sum per pass.
constexpr int kQagsWaveBlocks = 1024;  // workgroups of the overflow kernel (36.5 KB LDS each)

template <class TS>
__global__ __launch_bounds__(64) void k_flux_qags_wave(TS ts, const long long* __restrict__ overflow,
                                                       unsigned long long* __restrict__ stats,
                                                       double* __restrict__ flux, int* __restrict__ ier_out)
{
  __shared__ QagsBigArrays ws;
  __shared__ double vals[2][21];
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
        const QK21Out o = qk21_stored(vals[0], 1, iv[0][0], iv[0][1]);
        QK21Out o2{};
        if (np == 2) o2 = qk21_stored(vals[1], 1, iv[1][0], iv[1][1]);
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
        ts.failed(ft.slot);
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
  const unsigned char* need = nullptr;  // work items to run (null: all)
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
  static constexpr int kWavesPerEu = 5;  // register budget of k_lumi_fp64 (waves per SIMD)
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
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int q = p0 + wv;
  const bool live = q < lumi_work_items(v.npts, v.fp.mirror_ny, v.mirror_out) && (!v.need || v.need[q]);
  // a masked launch (block-uniform v.need): blocks with no needed item end here
  if (v.need && !__syncthreads_or(live)) return;
  for (int i = threadIdx.x; i < kNB - 1; i += blockDim.x) sm.gaa[i] = v.gaa[i];
  if (threadIdx.x < 5) {
    const double c = cos(kPi * kGLx5[threadIdx.x]);
    const double sn = sin(kPi * kGLx5[threadIdx.x]);
    sm.cosk[threadIdx.x] = c;
    sm.w5[threadIdx.x] = kGLw5[threadIdx.x];
    sm.w5[5 + threadIdx.x] = sn * sn;
  }
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
// JR::kWavesPerEu budgets the registers: the breakup-spline gathers from L2
// bound the kernel, and for one grid a fifth wave per SIMD hides more of
// their latency (96 VGPRs, a few spilled; the LDS of 4-wave blocks allows 5
// blocks per CU: 1.58 -> 1.48 ms on the 1001 x 121 grid), while the batch's
// mostly polarised jobs spill more at that budget and keep 4.
template <int WAVES, class JR>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(JR::kWavesPerEu))) void k_lumi_fp64(
    JR jr)
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
// The FP64 contraction of a view (only its work items with v.need[q] != 0
// when v.need is set) on stream s.
int launch_lumi_fp64(const LumiView& v, hipStream_t s);
// Copies the QAGS counters of the last prepare_flux into the context.
int read_flux_stats(upc_ctx* c, hipStream_t s);

}  // namespace upc
