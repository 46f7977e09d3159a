// Photon-flux tables and the FP64 two-photon luminosity on the GPU.
//
//   fluxPoint / fluxFormIntegrand / fluxForm   src/UpcCrossSection.cpp:166-218
//   calcTwoPhotonLumi / calcTwoPhotonLumiPol    src/UpcCrossSection.cpp:221-335
//   prepareTwoPhotonLumi grid loop              src/UpcCrossSection.cpp:494-576
//
// Layout: a flux row is the 120 values N(b_j, k) on the log b grid of photon
// energy k.  Point p (of npts) uses row p for k1 = M/2 e^Y (g1) and a second
// row for k2 = M/2 e^-Y (g2).  On a grid whose Y range is symmetric (ymin =
// -ymax, lower bin edges) the k2 of bin (im, iy >= 1) is the k1 of its mirror
// bin (im, ny - iy) up to one rounding of y, so only the iy = 0 column gets
// own k2 rows: rows [npts, npts + nm).  Otherwise rows [npts, 2 npts) hold the
// k2 rows.  flux[r*120 + j].  Nodes with b <= 2R (form-factor flux) become
// QAGS tasks; the rest take the point-charge flux.
#include "upc_internal.h"
#include "upc_qags.h"
#include "upc_flux.h"

#include <climits>

namespace upc {

constexpr int kQagsCap = 16;       // fast per-lane interval list (the flux integrals stop by 15)
constexpr int kQagsLimit = 1000;   // gsl_integration_qags limit of the reference
constexpr int kQagsBigCap = 1000;

__device__ __forceinline__ double flux_point(double b, double k, double g1, double factor)
{
  const double g = g1;
  const double x = b * k / g / kHc;
  double K0 = 0, K1 = 0;
  if (x > 1e-10) bessel_k01(x, &K0, &K1);
  return factor * k / g / g * (K1 * K1 + K0 * K0 / g / g);
}

// fluxFormIntegrand: q^2 F(min(t, Q2max - dQ2)) / t * J1(b q / hc)
struct FluxFormIntegrand {
  const double4* __restrict__ ff;
  double b, w2g2;
  // The spline record is a dependent load from a 32 MB table: it is issued
  // first and J1 (independent ALU work) runs while it is in flight.
  __device__ __forceinline__ double operator()(double q) const
  {
    const double t = q * q + w2g2;
    const double tc = t < kQ2max ? t : kQ2max - kDQ2;
    double xl;
    const int i = spline_interval(kQ2min, kDQ2, 1.0 / kDQ2, kNQ2, tc, &xl);
    const double4 c = ff[i];
    const double j1 = bessel_j1(b * q * (1.0 / kHc));
    const double d = tc - xl;
    const double f = c.x + d * (c.y + d * (c.z + d * c.w));
    return q * q * f / t * j1;
  }
};

// ---------------------------------------------------------------------------
// Grid points: lower bin edges (src/UpcCrossSection.cpp:540-550)
// ---------------------------------------------------------------------------
__global__ void k_grid_points(int nm, double mmin, double mmax, int ny, double ymin, double ymax,
                              double* __restrict__ pts)
{
  const long bin = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (bin >= (long)nm * ny) return;
  const int im = (int)(bin / ny), iy = (int)(bin % ny);
  const double dy = (ymax - ymin) / ny;
  const double dm = (mmax - mmin) / nm;
  pts[4 * bin + 0] = mmin + dm * im;
  pts[4 * bin + 1] = ymin + dy * iy;
  pts[4 * bin + 2] = dm;
  pts[4 * bin + 3] = dy;
}

// Row geometry and the number of form-factor (QAGS) nodes of each row.
__global__ void k_rows(const double* __restrict__ pts, int npts, int nrows, FluxParams fp,
                       double* __restrict__ row_k, double* __restrict__ row_logd, int* __restrict__ row_cnt)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
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
  row_k[r] = k;
  row_logd[r] = logd;
  row_cnt[r] = cnt;
}

// Exclusive scan of the row counts (one block); writes the total to *total.
__global__ __launch_bounds__(1024) void k_scan_counts(const int* __restrict__ cnt, int n,
                                                      int* __restrict__ off, unsigned long long* total)
{
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int s = t * per, e = min(n, s + per);
  long long acc = 0;
  for (int i = s; i < e; ++i) acc += cnt[i];
  part[t] = acc;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const long long v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long base = t ? part[t - 1] : 0;
  for (int i = s; i < e; ++i) {
    off[i] = (int)base;
    base += cnt[i];
  }
  if (t == 1023) *total = (unsigned long long)part[1023];
}

// Point-charge flux for nodes with b > 2R (or FLUX_POINT 1); QAGS task ids
// for the others.
__global__ void k_flux_point_rows(int nrows, FluxParams fp, const double* __restrict__ row_k,
                                  const double* __restrict__ row_logd, const int* __restrict__ row_cnt,
                                  const int* __restrict__ row_off, double* __restrict__ flux,
                                  int* __restrict__ tasks)
{
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)nrows * kNB1) return;
  const int r = (int)(gid / kNB1), j = (int)(gid % kNB1);
  if (j < row_cnt[r]) {
    tasks[row_off[r] + j] = r * 128 + j;
    return;
  }
  double bl, bh;
  const double b = row_node(fp.bmin, row_logd[r], j, &bl, &bh);
  flux[gid] = flux_point(b, row_k[r], fp.g1, fp.factor);
}

// ---------------------------------------------------------------------------
// Form-factor flux: persistent QAGS kernel.  Each lane owns one integral at a
// time; lanes take new tasks from a global cursor as they finish, so the
// wavefront keeps issuing full-width QK21 passes.
// ---------------------------------------------------------------------------
struct GridTask {
  const int* __restrict__ tasks;
  const double* __restrict__ row_k;
  const double* __restrict__ row_logd;
  double bmin;
  __device__ void get(int t, double* b, double* k, long* out) const
  {
    const int id = tasks[t];
    const int r = id >> 7, j = id & 127;
    double bl, bh;
    *b = row_node(bmin, row_logd[r], j, &bl, &bh);
    *k = row_k[r];
    *out = (long)r * kNB1 + j;
  }
};

struct PointTask {
  const double* __restrict__ b;
  const double* __restrict__ k;
  const int* __restrict__ idx;  // compacted indices of points needing QAGS
  __device__ void get(int t, double* bb, double* kk, long* out) const
  {
    const int i = idx[t];
    *bb = b[i];
    *kk = k[i];
    *out = i;
  }
};

template <int CAP, class TS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_flux_qags(TS ts, const unsigned long long* __restrict__ ntasks_p,
                                                   unsigned long long* __restrict__ cursor,
                                                   FluxParams fp, const double4* __restrict__ ff,
                                                   double* __restrict__ flux, int* __restrict__ ier_out,
                                                   unsigned long long* __restrict__ stats,
                                                   int* __restrict__ overflow)
{
  __shared__ double fvbuf[20 * 256];
  Qags<CAP> q;
  const long ntasks = (long)*ntasks_p;
  const int lane = threadIdx.x & 63;
  double* fv = fvbuf + (threadIdx.x >> 6) * 20 * 64 + lane;
  long task = -1;      // -1: needs a task; LONG_MAX: exhausted
  double k = 0;
  long out = 0;
  FluxFormIntegrand f;
  f.ff = ff;
  unsigned long long my_evals = 0, my_bad = 0;
  while (true) {
    const bool need = task < 0;
    const unsigned long long mask = __ballot(need);
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(cursor, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        const unsigned long long rank = __popcll(mask & ((1ull << lane) - 1ull));
        const long t = (long)(base + rank);
        if (t < ntasks) {
          task = t;
          double b;
          ts.get((int)t, &b, &k, &out);
          f.b = b;
          f.w2g2 = k * k / fp.g1 / fp.g1;
          q.start(0., 10., 1e-4, 1e-4, kQagsLimit);
        } else {
          task = LONG_MAX;
        }
      }
    }
    if (__ballot(task != LONG_MAX) == 0) break;
    if (task != LONG_MAX) {
      double qa, qb;
      q.interval(&qa, &qb);
      const QK21Out o = qk21_lds(f, qa, qb, fv, 64);
      const int st = q.consume(o);
      if (st == kQagsDone) {
        const double Q = q.result / fp.A;
        flux[out] = fp.factor * Q * Q / k;
        if (ier_out) ier_out[out] = q.ier;
        my_evals += (unsigned long long)q.neval();
        my_bad += q.ier != 0;
        task = -1;
      } else if (st == kQagsOverflow) {
        const unsigned long long slot = atomicAdd(&stats[2], 1ull);
        overflow[slot] = (int)task;
        task = -1;
      }
    }
  }
  // per-wave aggregation of statistics
  for (int o = 32; o > 0; o >>= 1) {
    my_evals += __shfl_down(my_evals, o);
    my_bad += __shfl_down(my_bad, o);
  }
  if (lane == 0) {
    atomicAdd(&stats[0], my_evals);
    atomicAdd(&stats[1], my_bad);
  }
}

// Tasks that outgrew the fast interval list: rerun with the reference's
// 1000-interval workspace, one per lane, workspace in global memory.
template <class TS>
__global__ __launch_bounds__(64) void k_flux_qags_big(TS ts, const int* __restrict__ overflow,
                                                      unsigned long long* __restrict__ stats,
                                                      FluxParams fp, const double4* __restrict__ ff,
                                                      double* __restrict__ flux, int* __restrict__ ier_out,
                                                      Qags<kQagsBigCap>* __restrict__ ws, int nws)
{
  const long n = (long)stats[2];
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nws) return;
  __shared__ double fvbuf[20 * 64];
  double* fv = fvbuf + threadIdx.x;
  Qags<kQagsBigCap>& q = ws[gid];
  FluxFormIntegrand f;
  f.ff = ff;
  for (long i = gid; i < n; i += nws) {
    double b, k;
    long out;
    ts.get(overflow[i], &b, &k, &out);
    f.b = b;
    f.w2g2 = k * k / fp.g1 / fp.g1;
    q.start(0., 10., 1e-4, 1e-4, kQagsLimit);
    int st = kQagsRunning;
    while (st == kQagsRunning) {
      double qa, qb;
      q.interval(&qa, &qb);
      st = q.consume(qk21_lds(f, qa, qb, fv, 64));
    }
    const double Q = q.result / fp.A;
    flux[out] = fp.factor * Q * Q / k;
    if (ier_out) ier_out[out] = st == kQagsDone ? q.ier : 99;
    atomicAdd(&stats[0], (unsigned long long)q.neval());
    if (st != kQagsDone || q.ier != 0) atomicAdd(&stats[1], 1ull);
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
constexpr int kLumiWaves = 4;

template <bool POL>
__global__ __launch_bounds__(64 * kLumiWaves) void k_lumi_fp64(
    const double* __restrict__ pts, int npts, const double* __restrict__ row_logd,
    const double* __restrict__ flux, FluxParams fp, int mode, const double4* __restrict__ gaa_g,
    const double4* __restrict__ bk, const DevScalars* __restrict__ sc, double* __restrict__ L,
    double* __restrict__ Ls, double* __restrict__ Lp)
{
  __shared__ double4 gaa[kNB];
  __shared__ double sb1[kLumiWaves][kNB1], sb2[kLumiWaves][kNB2];
  __shared__ double sw1[kLumiWaves][kNB1], sw2[kLumiWaves][kNB2];
  __shared__ int spre[kLumiWaves][kNB1 + 1], sjlo[kLumiWaves][kNB1];
  __shared__ double cosk[5], w5[10];
  for (int i = threadIdx.x; i < kNB - 1; i += blockDim.x) gaa[i] = gaa_g[i];
  if (threadIdx.x < 5) {
    const double c = cos(kPi * kGLx5[threadIdx.x]);
    const double sn = sin(kPi * kGLx5[threadIdx.x]);
    cosk[threadIdx.x] = c;
    w5[threadIdx.x] = kGLw5[threadIdx.x];
    w5[5 + threadIdx.x] = sn * sn;
  }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int p = blockIdx.x * kLumiWaves + wv;
  const bool live = p < npts;
  const double p20 = sc->p20;
  double M = 0, dm = 1, dy = 1;
  if (live) {
    M = pts[4 * p];
    dm = pts[4 * p + 2];
    dy = pts[4 * p + 3];
    const int r2 = k2_row(p, npts, fp.mirror_ny);
    const double logd1 = row_logd[p], logd2 = row_logd[r2];
    const double* f1 = flux + (long)p * kNB1;
    const double* f2 = flux + (long)r2 * kNB1;
    for (int j = lane; j < kNB1; j += 64) {
      double bl, bh;
      const double b1 = row_node(fp.bmin, logd1, j, &bl, &bh);
      sb1[wv][j] = b1;
      sw1[wv][j] = f1[j] * b1 * (bh - bl);
      const double b2 = row_node(fp.bmin, logd2, j, &bl, &bh);
      sb2[wv][j] = b2;
      sw2[wv][j] = f2[j] * b2 * (bh - bl);
    }
  }
  __syncthreads();
  // the extreme-cos node has the smallest b for every pair
  const double cext = POL ? -cosk[0] : cosk[4];
  if (live) {
    for (int i = lane; i < kNB1; i += 64) {
      const double b1 = sb1[wv][i];
      int lo = -1, hi = -1;
      for (int j = 0; j < kNB2; ++j) {
        if (node_b2(b1, sb2[wv][j], cext) < 400.0) {
          if (lo < 0) lo = j;
          hi = j;
        }
      }
      sjlo[wv][i] = lo < 0 ? 0 : lo;
      spre[wv][i + 1] = lo < 0 ? 0 : hi - lo + 1;
    }
  }
  __syncthreads();
  if (live && lane == 0) {
    spre[wv][0] = 0;
    for (int i = 1; i <= kNB1; ++i) spre[wv][i] += spre[wv][i - 1];
  }
  __syncthreads();
  if (!live) return;
  // S0: the phi sum when every node is at or beyond 20 fm
  double s0s = 0, s0p = 0;
  {
    const double br = mode != 1 ? p20 : 1.;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (POL) {
        s0s += br * 1. * w5[k] * cosk[k] * cosk[k];
        s0p += br * 1. * w5[k] * w5[5 + k];
      } else {
        s0s += br * 1. * w5[k];
      }
    }
  }
  const int T = spre[wv][kNB1];
  double accs = 0, accp = 0;
  for (int t = lane; t < T; t += 64) {
    int lo = 0, hi = kNB1;  // largest i with pre[i] <= t
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (spre[wv][mid] <= t) lo = mid;
      else hi = mid;
    }
    const int i = lo;
    const int j = sjlo[wv][i] + (t - spre[wv][i]);
    double ks, kp;
    node_k<POL>(sb1[wv][i], sb2[wv][j], cosk, w5, mode, p20, gaa, bk, &ks, &kp);
    const double ww = sw1[wv][i] * sw2[wv][j];
    accs += ww * (ks - s0s);
    if (POL) accp += ww * (kp - s0p);
  }
  double sw1t = 0, sw2t = 0;
  for (int j = lane; j < kNB1; j += 64) {
    sw1t += sw1[wv][j];
    sw2t += sw2[wv][j];
  }
  for (int o = 32; o > 0; o >>= 1) {
    accs += __shfl_down(accs, o);
    accp += __shfl_down(accp, o);
    sw1t += __shfl_down(sw1t, o);
    sw2t += __shfl_down(sw2t, o);
  }
  if (lane == 0) {
    const double pre = 2 * kPi * kPi * M;
    const double tot_s = s0s * (sw1t * sw2t) + accs;
    if (POL) {
      const double tot_p = s0p * (sw1t * sw2t) + accp;
      Ls[p] = pre * tot_s * dm * dy;
      Lp[p] = pre * tot_p * dm * dy;
    } else {
      L[p] = pre * tot_s * dm * dy;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

static int qags_persistent_blocks()
{
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (ncu <= 0) ncu = 256;
  return ncu * 4;
}

static int ensure_work(upc_ctx* c, size_t npts)
{
  LumiWork& w = c->work;
  const size_t rows = 2 * npts;
  if (rows <= w.rows_cap && npts <= w.points_cap) return 0;
  free_work(c);
  UPC_HIP(hipMalloc(&w.d_row_k, sizeof(double) * rows));
  UPC_HIP(hipMalloc(&w.d_row_logd, sizeof(double) * rows));
  UPC_HIP(hipMalloc(&w.d_row_cnt, sizeof(int) * rows));
  UPC_HIP(hipMalloc(&w.d_row_off, sizeof(int) * rows));
  UPC_HIP(hipMalloc(&w.d_flux, sizeof(double) * rows * kNB1));
  UPC_HIP(hipMalloc(&w.d_tasks, sizeof(int) * rows * kNB1));
  UPC_HIP(hipMalloc(&w.d_overflow, sizeof(int) * rows * kNB1));
  UPC_HIP(hipMalloc(&w.d_counters, sizeof(unsigned long long) * 8));
  UPC_HIP(hipMalloc(&w.d_points, sizeof(double) * 4 * npts));
  w.rows_cap = rows;
  w.tasks_cap = rows * kNB1;
  w.points_cap = npts;
  return 0;
}

void free_work(upc_ctx* c)
{
  LumiWork& w = c->work;
  void* ptrs[] = {w.d_row_k, w.d_row_logd, w.d_row_cnt, w.d_row_off, w.d_flux, w.d_tasks,
                  w.d_overflow, w.d_counters, w.d_points};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  w = LumiWork{};
}

static Qags<kQagsBigCap>* g_big_ws = nullptr;  // lazily allocated, process lifetime
static int g_big_ws_n = 0;

static int big_ws(int n, Qags<kQagsBigCap>** out)
{
  if (g_big_ws_n < n) {
    if (g_big_ws) (void)hipFree(g_big_ws);
    UPC_HIP(hipMalloc(&g_big_ws, sizeof(Qags<kQagsBigCap>) * n));
    g_big_ws_n = n;
  }
  *out = g_big_ws;
  return 0;
}

static int grid_mirror_ny(const upc_ctx* c, bool grid)
{
  return grid && c->cfg.ymin == -c->cfg.ymax ? c->cfg.ny : 0;
}

static int flux_rows(int npts, int mirror_ny) { return mirror_ny > 0 ? npts + npts / mirror_ny : 2 * npts; }

// Runs flux rows + QAGS for the points already in c->work.d_points.
static int run_flux(upc_ctx* c, int npts, int mirror_ny, hipStream_t s)
{
  LumiWork& w = c->work;
  const FluxParams fp = flux_params(c, mirror_ny);
  const int nrows = flux_rows(npts, mirror_ny);
  {
  PhaseScope ph(c, s, UPC_PHASE_FLUX);
  UPC_HIP(hipMemsetAsync(w.d_counters, 0, sizeof(unsigned long long) * 8, s));
  hipLaunchKernelGGL(k_rows, dim3(nblk(nrows, 256)), dim3(256), 0, s, w.d_points, npts, nrows, fp, w.d_row_k,
                     w.d_row_logd, w.d_row_cnt);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, w.d_row_cnt, nrows, w.d_row_off,
                     w.d_counters + 4);
  hipLaunchKernelGGL(k_flux_point_rows, dim3(nblk((long)nrows * kNB1, 256)), dim3(256), 0, s, nrows, fp,
                     w.d_row_k, w.d_row_logd, w.d_row_cnt, w.d_row_off, w.d_flux, w.d_tasks);
  UPC_HIP(hipGetLastError());
  }
  if (!c->cfg.flux_point) {
    PhaseScope ph(c, s, UPC_PHASE_QAGS);
    GridTask ts{w.d_tasks, w.d_row_k, w.d_row_logd, fp.bmin};
    hipLaunchKernelGGL((k_flux_qags<kQagsCap, GridTask>), dim3(qags_persistent_blocks()), dim3(256), 0, s, ts,
                       w.d_counters + 4, w.d_counters + 5, fp, c->d_ff, w.d_flux, (int*)nullptr,
                       w.d_counters, w.d_overflow);
    Qags<kQagsBigCap>* ws = nullptr;
    int rc = big_ws(256, &ws);
    if (rc) return rc;
    hipLaunchKernelGGL((k_flux_qags_big<GridTask>), dim3(4), dim3(64), 0, s, ts, w.d_overflow, w.d_counters,
                       fp, c->d_ff, w.d_flux, (int*)nullptr, ws, 256);
    UPC_HIP(hipGetLastError());
  }
  return 0;
}

static int read_stats(upc_ctx* c, hipStream_t s)
{
  unsigned long long h[8];
  UPC_HIP(hipMemcpyAsync(h, c->work.d_counters, sizeof(h), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipStreamSynchronize(s));
  c->last_qags_evals = (long long)h[0];
  c->last_qags_bad = (long long)h[1];
  c->last_qags_calls = (long long)h[4];
  return 0;
}

int make_grid_points(upc_ctx* c, hipStream_t s)
{
  const upc_config& g = c->cfg;
  const long npts = (long)g.nm * g.ny;
  int rc = ensure_work(c, (size_t)npts);
  if (rc) return rc;
  hipLaunchKernelGGL(k_grid_points, dim3(nblk(npts, 256)), dim3(256), 0, s, g.nm, g.mmin, g.mmax, g.ny,
                     g.ymin, g.ymax, c->work.d_points);
  UPC_HIP(hipGetLastError());
  return 0;
}

__global__ void k_copy_points(const double* __restrict__ M, const double* __restrict__ Y, int n,
                              double* __restrict__ pts)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pts[4 * i] = M[i];
  pts[4 * i + 1] = Y[i];
  pts[4 * i + 2] = 1.0;
  pts[4 * i + 3] = 1.0;
}

int prepare_flux(upc_ctx* c, const double* d_pts, int npts, FluxParams* fp, hipStream_t s)
{
  int rc = ensure_work(c, (size_t)npts);
  if (rc) return rc;
  if (d_pts) {
    hipLaunchKernelGGL(k_copy_points, dim3(nblk(npts, 256)), dim3(256), 0, s, d_pts, d_pts + npts, npts,
                       c->work.d_points);
  }
  const int mirror_ny = grid_mirror_ny(c, d_pts == nullptr);
  rc = run_flux(c, npts, mirror_ny, s);
  if (rc) return rc;
  *fp = flux_params(c, mirror_ny);
  return 0;
}

int read_flux_stats(upc_ctx* c, hipStream_t s) { return read_stats(c, s); }

// d_pts == nullptr: the grid points already in the work buffer.
int lumi_points_fp64(upc_ctx* c, const double* d_pts, int npts, double* d_L, double* d_Ls, double* d_Lp,
                     hipStream_t s)
{
  if (npts <= 0) return 0;
  FluxParams fp;
  int rc = prepare_flux(c, d_pts, npts, &fp, s);
  if (rc) return rc;
  const dim3 grid(nblk(npts, kLumiWaves)), block(64 * kLumiWaves);
  {
  PhaseScope ph(c, s, UPC_PHASE_LUMI);
  if (c->cfg.use_polarized) {
    hipLaunchKernelGGL(k_lumi_fp64<true>, grid, block, 0, s, c->work.d_points, npts, c->work.d_row_logd,
                       c->work.d_flux, fp, c->cfg.breakup_mode, c->d_gaa, c->d_bk, c->d_scalars,
                       (double*)nullptr, d_Ls, d_Lp);
  } else {
    hipLaunchKernelGGL(k_lumi_fp64<false>, grid, block, 0, s, c->work.d_points, npts, c->work.d_row_logd,
                       c->work.d_flux, fp, c->cfg.breakup_mode, c->d_gaa, c->d_bk, c->d_scalars, d_L,
                       (double*)nullptr, (double*)nullptr);
  }
  }
  UPC_HIP(hipGetLastError());
  return read_stats(c, s);
}

// ---- batched per-point flux (C-ABI upc_flux_point / upc_flux_form) ----
__global__ void k_flux_point_pts(const double* __restrict__ b, const double* __restrict__ k, int n,
                                 FluxParams fp, double* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = flux_point(b[i], k[i], fp.g1, fp.factor);
}

// fluxForm: point flux where isPoint or b > 2R, else a QAGS task.
__global__ void k_flux_form_classify(const double* __restrict__ b, const double* __restrict__ k, int n,
                                     FluxParams fp, double* __restrict__ out, int* __restrict__ ier,
                                     int* __restrict__ idx, unsigned long long* __restrict__ counters)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fp.isPoint || b[i] > fp.twoR) {
    out[i] = flux_point(b[i], k[i], fp.g1, fp.factor);
    if (ier) ier[i] = 0;
    return;
  }
  const unsigned long long slot = atomicAdd(&counters[4], 1ull);
  idx[slot] = i;
}

int launch_flux_point(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_flux_point_pts, dim3(nblk(n, 256)), dim3(256), 0, s, d_b, d_k, n, flux_params(c), d_out);
  UPC_HIP(hipGetLastError());
  return 0;
}

int launch_flux_form(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out, int* d_ier,
                     hipStream_t s)
{
  if (n <= 0) return 0;
  int rc = ensure_work(c, (size_t)(n + 1) / 2 + 1);
  if (rc) return rc;
  LumiWork& w = c->work;
  const FluxParams fp = flux_params(c);
  UPC_HIP(hipMemsetAsync(w.d_counters, 0, sizeof(unsigned long long) * 8, s));
  hipLaunchKernelGGL(k_flux_form_classify, dim3(nblk(n, 256)), dim3(256), 0, s, d_b, d_k, n, fp, d_out, d_ier,
                     w.d_tasks, w.d_counters);
  PointTask ts{d_b, d_k, w.d_tasks};
  hipLaunchKernelGGL((k_flux_qags<kQagsCap, PointTask>), dim3(qags_persistent_blocks()), dim3(256), 0, s, ts,
                     w.d_counters + 4, w.d_counters + 5, fp, c->d_ff, d_out, d_ier, w.d_counters, w.d_overflow);
  Qags<kQagsBigCap>* ws = nullptr;
  rc = big_ws(256, &ws);
  if (rc) return rc;
  hipLaunchKernelGGL((k_flux_qags_big<PointTask>), dim3(4), dim3(64), 0, s, ts, w.d_overflow, w.d_counters, fp,
                     c->d_ff, d_out, d_ier, ws, 256);
  UPC_HIP(hipGetLastError());
  return read_stats(c, s);
}

}  // namespace upc
