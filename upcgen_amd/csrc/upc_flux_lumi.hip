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
  row_setup(pts, npts, r, fp, &row_k[r], &row_logd[r], &row_cnt[r]);
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
  point_row_node((int)(gid / kNB1), (int)(gid % kNB1), fp, row_k, row_logd, row_cnt, row_off, flux, tasks);
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
    GridTask ts{w.d_tasks, w.d_row_k, w.d_row_logd, fp, c->d_ff};
    hipLaunchKernelGGL((k_flux_qags<kQagsCap, GridTask>), dim3(qags_persistent_blocks()), dim3(256), 0, s, ts,
                       w.d_counters + 4, w.d_counters + 5, w.d_flux, (int*)nullptr, w.d_counters,
                       w.d_overflow);
    Qags<kQagsBigCap>* ws = nullptr;
    int rc = big_ws(256, &ws);
    if (rc) return rc;
    hipLaunchKernelGGL((k_flux_qags_big<GridTask>), dim3(4), dim3(64), 0, s, ts, w.d_overflow, w.d_counters,
                       w.d_flux, (int*)nullptr, ws, 256);
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
  constexpr int kWaves = 4;
  SingleLumi jr{LumiView{c->work.d_points, c->work.d_row_logd, c->work.d_flux, fp, c->d_gaa, c->d_bk,
                         c->d_scalars, npts, c->cfg.breakup_mode, c->cfg.use_polarized, d_L, d_Ls, d_Lp}};
  {
    PhaseScope ph(c, s, UPC_PHASE_LUMI);
    hipLaunchKernelGGL((k_lumi_fp64<kWaves, SingleLumi>), dim3(nblk(npts, kWaves)), dim3(64 * kWaves), 0, s, jr);
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
  PointTask ts{d_b, d_k, w.d_tasks, fp, c->d_ff};
  hipLaunchKernelGGL((k_flux_qags<kQagsCap, PointTask>), dim3(qags_persistent_blocks()), dim3(256), 0, s, ts,
                     w.d_counters + 4, w.d_counters + 5, d_out, d_ier, w.d_counters, w.d_overflow);
  Qags<kQagsBigCap>* ws = nullptr;
  rc = big_ws(256, &ws);
  if (rc) return rc;
  hipLaunchKernelGGL((k_flux_qags_big<PointTask>), dim3(4), dim3(64), 0, s, ts, w.d_overflow, w.d_counters,
                     d_out, d_ier, ws, 256);
  UPC_HIP(hipGetLastError());
  return read_stats(c, s);
}

}  // namespace upc
