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

#include <algorithm>
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
// need (may be null): rows with need[r] == 0 get no QAGS tasks (their
// form-factor nodes are left unset).
__global__ void k_rows(const double* __restrict__ pts, int npts, int nrows, FluxParams fp,
                       double* __restrict__ row_k, double* __restrict__ row_logd, int* __restrict__ row_cnt,
                       const unsigned char* __restrict__ need)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  row_setup(pts, npts, r, fp, &row_k[r], &row_logd[r], &row_cnt[r]);
  if (need && !need[r]) row_cnt[r] = 0;
}

// Point-charge flux for nodes with b > 2R (or FLUX_POINT 1); QAGS task ids
// for the others.
// row_need (may be null): rows that no needed bin reads are skipped.
__global__ void k_flux_point_rows(int nrows, FluxParams fp, const double* __restrict__ row_k,
                                  const double* __restrict__ row_logd, const int* __restrict__ row_cnt,
                                  const long long* __restrict__ row_off, double* __restrict__ flux,
                                  TaskRec* __restrict__ recs, const unsigned char* __restrict__ row_need)
{
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)nrows * kNB1) return;
  const long r = gid / kNB1;
  if (row_need && !row_need[r]) return;
  point_row_node(r, (int)(gid % kNB1), fp, row_k, row_logd, row_cnt, row_off, flux, recs);
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
  return ncu * kQagsBlocksPerCu;  // as many workgroups as stay resident
}

// QAGS nodes of one flux row are at most 97: node j has b_j >= bmin e^{j logd}
// with bmin = 0.05 R and logd >= ln(5R / bmin) / 120, so b_j <= 2R needs
// j <= 120 ln 40 / ln 100 = 96.09.
constexpr int kMaxQagsNodesPerRow = 98;

// Work buffers for npts points and their flux rows (at most 2 npts); task
// records, offsets and overflow lists are 64-bit.  The point-charge flux has
// no QAGS tasks, so its task buffers stay minimal; the per-point form-factor
// call needs `tasks` of them (its point indices).
static int ensure_work(upc_ctx* c, size_t npts, size_t rows, size_t tasks = 0)
{
  LumiWork& w = c->work;
  const size_t need_tasks = std::max<size_t>(
      {tasks, c->cfg.flux_point ? (size_t)1 : rows * kMaxQagsNodesPerRow, (size_t)1});
  if (rows <= w.rows_cap && npts <= w.points_cap && need_tasks <= w.tasks_cap) return 0;
  rows = std::max(rows, w.rows_cap);
  npts = std::max(npts, w.points_cap);
  const size_t ntask = std::max(need_tasks, w.tasks_cap);
  free_work(c);
  UPC_HIP(hipMalloc(&w.d_row_k, sizeof(double) * rows));
  UPC_HIP(hipMalloc(&w.d_row_logd, sizeof(double) * rows));
  UPC_HIP(hipMalloc(&w.d_row_cnt, sizeof(int) * rows));
  UPC_HIP(hipMalloc(&w.d_row_off, sizeof(long long) * rows));
  UPC_HIP(hipMalloc(&w.d_flux, sizeof(double) * rows * kNB1));
  UPC_HIP(hipMalloc(&w.d_tasks, sizeof(TaskRec) * ntask));
  UPC_HIP(hipMalloc(&w.d_overflow, sizeof(long long) * ntask));
  UPC_HIP(hipMalloc(&w.d_counters, sizeof(unsigned long long) * 8));
  UPC_HIP(hipMalloc(&w.d_points, sizeof(double) * 4 * npts));
  UPC_HIP(hipMalloc(&w.d_part, sizeof(long long) * ((rows + 1023) / 1024 + 1)));
  w.rows_cap = rows;
  w.tasks_cap = ntask;
  w.points_cap = npts;
  return 0;
}

void free_work(upc_ctx* c)
{
  LumiWork& w = c->work;
  void* ptrs[] = {w.d_row_k, w.d_row_logd, w.d_row_cnt, w.d_row_off, w.d_flux,
                  w.d_tasks, w.d_overflow, w.d_counters, w.d_points, w.d_part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  w = LumiWork{};
}

static int grid_mirror_ny(const upc_ctx* c, bool grid)
{
  return grid && c->cfg.ymin == -c->cfg.ymax ? c->cfg.ny : 0;
}

static long long flux_rows(int npts, int mirror_ny)
{
  return mirror_ny > 0 ? (long long)npts + npts / mirror_ny : 2LL * npts;
}

// Runs flux rows + QAGS for the points already in c->work.d_points.
static int run_flux(upc_ctx* c, int npts, int mirror_ny, hipStream_t s, const unsigned char* row_need = nullptr)
{
  LumiWork& w = c->work;
  const FluxParams fp = flux_params(c, mirror_ny);
  const int nrows = flux_rows(npts, mirror_ny);
  {
  PhaseScope ph(c, s, UPC_PHASE_FLUX);
  UPC_HIP(hipMemsetAsync(w.d_counters, 0, sizeof(unsigned long long) * 8, s));
  hipLaunchKernelGGL(k_rows, dim3(nblk(nrows, 256)), dim3(256), 0, s, w.d_points, npts, nrows, fp, w.d_row_k,
                     w.d_row_logd, w.d_row_cnt, row_need);
  const int nparts = nblk(nrows, kScanBlock);
  hipLaunchKernelGGL(k_scan_part<>, dim3(nparts), dim3(kScanBlock), 0, s, w.d_row_cnt, (long)nrows, w.d_part);
  hipLaunchKernelGGL(k_scan_top<>, dim3(1), dim3(kScanBlock), 0, s, w.d_part, nparts, w.d_counters + 4);
  hipLaunchKernelGGL(k_scan_apply<>, dim3(nparts), dim3(kScanBlock), 0, s, w.d_row_cnt, (long)nrows, w.d_part,
                     w.d_row_off);
  hipLaunchKernelGGL(k_flux_point_rows, dim3(nblk((long)nrows * kNB1, 256)), dim3(256), 0, s, nrows, fp,
                     w.d_row_k, w.d_row_logd, w.d_row_cnt, w.d_row_off, w.d_flux, (TaskRec*)w.d_tasks, row_need);
  UPC_HIP(hipGetLastError());
  }
  if (!c->cfg.flux_point) {
    PhaseScope ph(c, s, UPC_PHASE_QAGS);
    GridTask ts{(const TaskRec*)w.d_tasks, fp, c->d_ff};
    hipLaunchKernelGGL((k_flux_qags<kQagsCap, GridTask>), dim3(qags_persistent_blocks()), dim3(kQagsBlock), 0, s,
                       ts, w.d_counters + 4, w.d_counters + 5, w.d_flux, (int*)nullptr, w.d_counters,
                       w.d_overflow, c->qags_cap);
    hipLaunchKernelGGL((k_flux_qags_wave<GridTask>), dim3(kQagsWaveBlocks), dim3(64), 0, s, ts, w.d_overflow,
                       w.d_counters, w.d_flux, (int*)nullptr);
    UPC_HIP(hipGetLastError());
  }
  return 0;
}

static int read_stats(upc_ctx* c, hipStream_t s)
{
  UPC_HIP(hipMemcpyAsync(c->hp_stats, c->work.d_counters, sizeof(unsigned long long) * 8, hipMemcpyDeviceToHost, s));
  UPC_HIP(hipEventRecord(c->ev_stats, s));
  c->stats_pending = true;
  return 0;
}

int make_grid_points(upc_ctx* c, hipStream_t s)
{
  const upc_config& g = c->cfg;
  const long npts = (long)g.nm * g.ny;
  int rc = ensure_work(c, (size_t)npts, (size_t)flux_rows((int)npts, grid_mirror_ny(c, true)));
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
  const int mirror_ny = grid_mirror_ny(c, d_pts == nullptr);
  int rc = ensure_work(c, (size_t)npts, (size_t)flux_rows(npts, mirror_ny));
  if (rc) return rc;
  if (d_pts) {
    hipLaunchKernelGGL(k_copy_points, dim3(nblk(npts, 256)), dim3(256), 0, s, d_pts, d_pts + npts, npts,
                       c->work.d_points);
  }
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
  const int mirror_out = fp.mirror_ny > 0;
  SingleLumi jr{LumiView{c->work.d_points, c->work.d_row_logd, c->work.d_flux, fp, c->d_gaa, c->d_bk,
                         c->d_scalars, npts, c->cfg.breakup_mode, c->cfg.use_polarized, d_L, d_Ls, d_Lp,
                         mirror_out}};
  const int nwork = lumi_work_items(npts, fp.mirror_ny, mirror_out);
  {
    PhaseScope ph(c, s, UPC_PHASE_LUMI);
    hipLaunchKernelGGL((k_lumi_fp64<kWaves, SingleLumi>), dim3(nblk(nwork, kWaves)), dim3(64 * kWaves), 0, s,
                       jr);
  }
  UPC_HIP(hipGetLastError());
  return read_stats(c, s);
}

int launch_lumi_fp64(const LumiView& v, hipStream_t s)
{
  constexpr int kWaves = 4;
  const int nwork = lumi_work_items(v.npts, v.fp.mirror_ny, v.mirror_out);
  if (nwork <= 0) return 0;
  hipLaunchKernelGGL((k_lumi_fp64<kWaves, SingleLumi>), dim3(nblk(nwork, kWaves)), dim3(64 * kWaves), 0, s,
                     SingleLumi{v});
  UPC_HIP(hipGetLastError());
  return 0;
}

// Row and work-item masks of a grid restricted to the bins with
// need_bin != 0.  Row r < npts is bin r's k1 row and, on a symmetric grid,
// also the k2 row of its mirror bin (im, ny - iy) (iy >= 1); rows from npts
// on are k2 rows (of bin im * ny on a symmetric grid, of bin r - npts
// otherwise).  A work item is needed when either of its bins is.
__global__ void k_subset_masks(const unsigned char* __restrict__ need_bin, int npts, int nrows, int mirror_ny,
                               int nwork, int mirror_out, unsigned char* __restrict__ row_need,
                               unsigned char* __restrict__ item_need)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrows) {
    bool nd;
    if (i < npts) {
      nd = need_bin[i];
      if (mirror_ny > 0) {
        const int im = i / mirror_ny, iy = i - im * mirror_ny;
        if (iy > 0) nd = nd || need_bin[im * mirror_ny + (mirror_ny - iy)];
      }
    } else {
      nd = need_bin[mirror_ny > 0 ? (i - npts) * mirror_ny : i - npts];
    }
    row_need[i] = nd;
  }
  if (i < nwork) {
    int p, pm;
    work_point(i, mirror_ny, mirror_out, &p, &pm);
    item_need[i] = need_bin[p] || (pm >= 0 && need_bin[pm]);
  }
}

int lumi_grid_subset_fp64(upc_ctx* c, const unsigned char* d_need_bin, double* d_Lt, hipStream_t s)
{
  const upc_config& g = c->cfg;
  if (g.use_polarized) return set_error("the subset grid is unpolarised"), UPC_EINVAL;
  const int npts = g.nm * g.ny;
  int rc = make_grid_points(c, s);
  if (rc) return rc;
  const int mirror_ny = grid_mirror_ny(c, true);
  const int nrows = (int)flux_rows(npts, mirror_ny);
  const int mirror_out = mirror_ny > 0;
  const int nwork = lumi_work_items(npts, mirror_ny, mirror_out);
  unsigned char *row_need = nullptr, *item_need = nullptr;
  UPC_HIP(hipMallocAsync((void**)&row_need, (size_t)nrows, s));
  UPC_HIP(hipMallocAsync((void**)&item_need, (size_t)nwork, s));
  hipLaunchKernelGGL(k_subset_masks, dim3(nblk(std::max(nrows, nwork), 256)), dim3(256), 0, s, d_need_bin, npts,
                     nrows, mirror_ny, nwork, mirror_out, row_need, item_need);
  rc = run_flux(c, npts, mirror_ny, s, row_need);
  if (rc) return rc;
  const FluxParams fp = flux_params(c, mirror_ny);
  constexpr int kWaves = 4;
  LumiView v{c->work.d_points, c->work.d_row_logd, c->work.d_flux, fp, c->d_gaa, c->d_bk, c->d_scalars, npts,
             g.breakup_mode, 0, d_Lt, nullptr, nullptr, mirror_out};
  v.need = item_need;
  SingleLumi jr{v};
  {
    PhaseScope ph(c, s, UPC_PHASE_LUMI);
    hipLaunchKernelGGL((k_lumi_fp64<kWaves, SingleLumi>), dim3(nblk(nwork, kWaves)), dim3(64 * kWaves), 0, s,
                       jr);
  }
  (void)hipFreeAsync(row_need, s);
  (void)hipFreeAsync(item_need, s);
  UPC_HIP(hipGetLastError());
  return read_stats(c, s);
}

// ---- calcPhotonFlux / calcNucCrossSectionY (src/UpcCrossSection.cpp:699-748) ----
// One wave per point: sum_j P(min(b,20)) G(b) N(b_j, k) b_j (b_h - b_l) over
// the point's k1 flux row; flux = 2 pi k sum.
__global__ __launch_bounds__(256) void k_vm_flux(int n, const double* __restrict__ row_k,
                                                 const double* __restrict__ row_logd,
                                                 const double* __restrict__ flux, FluxParams fp, int mode,
                                                 const double4* __restrict__ gaa, const double4* __restrict__ bk,
                                                 double* __restrict__ out)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int p = blockIdx.x * 4 + wv;
  if (p >= n) return;
  const double db = 20. / (kNB - 1);
  double acc = 0;
  for (int j = lane; j < kNB1; j += 64) {
    double bl, bh;
    const double b = row_node(fp.bmin, row_logd[p], j, &bl, &bh);
    double breakup = 1.;
    if (mode != 1) breakup = spline_eval(bk, kBkMin, kDBk, 1.0 / kDBk, kNBkUsed, b < 20. ? b : 20.);
    const double g = b < 20. ? spline_eval(gaa, 0.0, db, 1.0 / db, kNB, b) : 1.;
    acc += breakup * g * flux[(long)p * kNB1 + j] * b * (bh - bl);
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) out[p] = 2. * kPi * row_k[p] * acc;
}

int launch_photon_flux(upc_ctx* c, const double* d_M, const double* d_Y, int n, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  double* pts = nullptr;
  UPC_HIP(hipMallocAsync((void**)&pts, sizeof(double) * 2 * n, s));
  UPC_HIP(hipMemcpyAsync(pts, d_M, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  UPC_HIP(hipMemcpyAsync(pts + n, d_Y, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  FluxParams fp;
  int rc = prepare_flux(c, pts, n, &fp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_vm_flux, dim3(nblk(n, 4)), dim3(256), 0, s, n, c->work.d_row_k, c->work.d_row_logd,
                     c->work.d_flux, fp, c->cfg.breakup_mode, c->d_gaa, c->d_bk, d_out);
  UPC_HIP(hipGetLastError());
  (void)hipFreeAsync(pts, s);
  return read_stats(c, s);
}

// cs[iy] = F(y) s(y) + F(-y) s(-y), ratio = first / second, total summed in
// iy order then x dy (single thread: the reference's order).
__global__ void k_nuc_cs_y(int ny, double dy, const double* __restrict__ f, const double* __restrict__ cs_pos,
                           const double* __restrict__ cs_neg, double* __restrict__ cs, double* __restrict__ ratio,
                           double* __restrict__ tot)
{
  double t = 0;
  for (int iy = 0; iy < ny; ++iy) {
    const double u1 = f[iy] * cs_pos[iy];
    const double u2 = f[ny + iy] * cs_neg[iy];
    cs[iy] = u1 + u2;
    t += u1 + u2;
    ratio[iy] = u1 / u2;
  }
  *tot = t * dy;
}

__global__ void k_vm_points(int ny, double ymin, double dy, double mpart, double* __restrict__ MY)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ny) return;
  const double y = ymin + dy * i;
  MY[i] = mpart;
  MY[ny + i] = mpart;
  MY[2 * ny + i] = y;
  MY[3 * ny + i] = -y;
}

int launch_nuc_cs_y(upc_ctx* c, double mpart, const double* d_cs_pos, const double* d_cs_neg, double* d_cs,
                    double* d_ratio, double* h_tot, hipStream_t s)
{
  const int ny = c->cfg.ny;
  const double dy = (c->cfg.ymax - c->cfg.ymin) / ny;
  double* buf = nullptr;  // M[2 ny], Y[2 ny], flux[2 ny], total
  UPC_HIP(hipMallocAsync((void**)&buf, sizeof(double) * (6 * ny + 1), s));
  hipLaunchKernelGGL(k_vm_points, dim3(nblk(ny, 256)), dim3(256), 0, s, ny, c->cfg.ymin, dy, mpart, buf);
  int rc = launch_photon_flux(c, buf, buf + 2 * ny, 2 * ny, buf + 4 * ny, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_nuc_cs_y, dim3(1), dim3(1), 0, s, ny, dy, buf + 4 * ny, d_cs_pos, d_cs_neg, d_cs, d_ratio,
                     buf + 6 * ny);
  double t = 0;
  UPC_HIP(hipMemcpyAsync(&t, buf + 6 * ny, sizeof(double), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipStreamSynchronize(s));
  *h_tot += t;
  (void)hipFreeAsync(buf, s);
  return 0;
}

// ---- J1 as the flux integrand evaluates it (C-ABI upc_bessel_j1) ----
__global__ void k_bessel_j1(const double* __restrict__ x, int n, double* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bessel_j1(x[i]);
}

int launch_bessel_j1(const double* d_x, int n, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_bessel_j1, dim3(nblk(n, 256)), dim3(256), 0, s, d_x, n, d_out);
  UPC_HIP(hipGetLastError());
  return 0;
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

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// The per-point form-factor flux on the buffers in w: classify (point flux
// beyond 2R, QAGS tasks inside), the persistent QAGS kernel, then the
// 1000-interval kernel for the tasks that outgrew the fast list.  Counters are cleared by a kernel, so the sequence may be captured.
int enqueue_flux_form(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out, int* d_ier,
                      const FluxFormBufs& w, hipStream_t s)
{
  const FluxParams fp = flux_params(c);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, w.counters, 8);
  hipLaunchKernelGGL(k_flux_form_classify, dim3(nblk(n, 256)), dim3(256), 0, s, d_b, d_k, n, fp, d_out, d_ier,
                     w.idx, w.counters);
  PointTask ts{d_b, d_k, w.idx, fp, c->d_ff};
  hipLaunchKernelGGL((k_flux_qags<kQagsCap, PointTask>), dim3(qags_persistent_blocks()), dim3(kQagsBlock), 0, s,
                     ts, w.counters + 4, w.counters + 5, d_out, d_ier, w.counters, w.overflow, c->qags_cap);
  hipLaunchKernelGGL((k_flux_qags_wave<PointTask>), dim3(kQagsWaveBlocks), dim3(64), 0, s, ts, w.overflow,
                     w.counters, d_out, d_ier);
  UPC_HIP(hipGetLastError());
  return 0;
}

int launch_flux_form(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out, int* d_ier,
                     hipStream_t s)
{
  if (n <= 0) return 0;
  // the per-point QAGS keeps its n point indices in the task buffer
  int rc = ensure_work(c, 1, 1, (size_t)n);
  if (rc) return rc;
  LumiWork& w = c->work;
  const FluxFormBufs bufs{(int*)w.d_tasks, w.d_overflow, w.d_counters};
  rc = enqueue_flux_form(c, d_b, d_k, n, d_out, d_ier, bufs, s);
  if (rc) return rc;
  return read_stats(c, s);
}

}  // namespace upc
