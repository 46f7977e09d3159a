// Batched luminosity grids: many collision systems, breakup modes and uneven
// (M, Y) grids in one launch sequence.
//
// Each job is one configuration, i.e. what one reference process computes in
// UpcCrossSection::init + prepareTwoPhotonLumi (src/UpcCrossSection.cpp:
// 116-137, 463-592).  Every job keeps its own context (tables: G_AA, form
// factor, breakup), and the grid work of all jobs is concatenated:
//   points [sum npts], flux rows [sum nrows][120], QAGS tasks [sum tasks].
// One kernel per stage serves all jobs (a job table in HBM maps a point,
// row or block to its job), so the QAGS integrals of every job share one
// persistent launch and its task cursor, and the contraction is one launch.
// The sequence is captured once into a hipGraph and replayed; buffers are
// cleared by kernels (no memset nodes).  The launch shapes (QAGS workgroups
// per CU, contraction waves per block) are chosen by timing the candidates
// on the batch itself.
#include "upc_flux.h"

#include <algorithm>
#include <climits>
#include <vector>

struct upc_batch;

namespace upc {

struct BatchJob {
  LumiView v;          // pts, row_logd, flux point at this job's slice
  const double4* ff;
  int nm, ny;
  double mmin, mmax, ymin, ymax;
  int pt0, row0;       // first point and row of the job in the batch arrays
};

// upc_mfma.hip: the MFMA contraction of one view, on a state of the batch's own
void* mfma_state_create(double h);
void mfma_state_free(void* st);
int mfma_state_ready(void* st, int prec, bool pol, size_t nrows);
int mfma_batch_contract(upc_ctx* c, void* st, const LumiView& v, int tile_group, hipStream_t s);
void mfma_zero_ksteps(void* st, hipStream_t s);
int mfma_read_counts(void* st, unsigned long long* h_out, hipStream_t s);

constexpr int kLumiWaveChoices[3] = {2, 4, 8};
constexpr int kMfmaTileChoices[3] = {1, 2, 4};  // column tiles per A-fragment load (MFMA form)
constexpr int kQagsBlockChoices[3] = {1, 2, 3};  // workgroups of 256 per CU (3 fit its LDS)

__device__ __forceinline__ int find_job(const int* __restrict__ start, int njobs, long idx)
{
  int lo = 0, hi = njobs - 1;  // largest j with start[j] <= idx
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= idx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_batch_zero(unsigned long long* __restrict__ p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_batch_points(const BatchJob* __restrict__ jobs, const int* __restrict__ pt0, int njobs,
                               long total, double* __restrict__ pts)
{
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const BatchJob& J = jobs[find_job(pt0, njobs, idx)];
  const long bin = idx - J.pt0;
  const int im = (int)(bin / J.ny), iy = (int)(bin % J.ny);
  const double dy = (J.ymax - J.ymin) / J.ny;
  const double dm = (J.mmax - J.mmin) / J.nm;
  pts[4 * idx + 0] = J.mmin + dm * im;
  pts[4 * idx + 1] = J.ymin + dy * iy;
  pts[4 * idx + 2] = dm;
  pts[4 * idx + 3] = dy;
}

__global__ void k_batch_rows(const BatchJob* __restrict__ jobs, const int* __restrict__ row0, int njobs, long total,
                             double* __restrict__ row_k, double* __restrict__ row_logd, int* __restrict__ row_cnt)
{
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= total) return;
  const BatchJob& J = jobs[find_job(row0, njobs, r)];
  row_setup(J.v.pts, J.v.npts, (int)(r - J.row0), J.v.fp, &row_k[r], &row_logd[r], &row_cnt[r]);
}

__global__ void k_batch_point_rows(const BatchJob* __restrict__ jobs, const int* __restrict__ row0, int njobs,
                                   long total_rows, const double* __restrict__ row_k,
                                   const double* __restrict__ row_logd, const int* __restrict__ row_cnt,
                                   const long long* __restrict__ row_off, double* __restrict__ flux,
                                   TaskRec* __restrict__ recs)
{
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total_rows * kNB1) return;
  const long r = gid / kNB1;
  const BatchJob& J = jobs[find_job(row0, njobs, r)];
  point_row_node(r, (int)(gid % kNB1), J.v.fp, row_k, row_logd, row_cnt, row_off, flux, recs);
}

// QAGS task source over all jobs: task records of the concatenated flux rows.
struct BatchTask {
  static constexpr bool kSharedFF = false;  // each job has its nucleus' table (FluxTask::ff)
  const TaskRec* __restrict__ recs;
  const BatchJob* __restrict__ jobs;
  const int* __restrict__ row0;
  int njobs;
  unsigned long long* __restrict__ job_failed;  // [njobs] integrals with ier != 0
  __device__ void get(long long t, FluxTask* ft) const
  {
    const TaskRec rec = recs[t];
    const int jb = find_job(row0, njobs, rec.out / kNB1);
    const BatchJob& J = jobs[jb];
    ft->b = rec.b;
    ft->k = rec.k;
    ft->out = rec.out;
    ft->ff = J.ff;
    ft->g1 = J.v.fp.g1;
    ft->factor = J.v.fp.factor;
    ft->A = J.v.fp.A;
    ft->slot = jb;
  }
  __device__ void out(long long t, int job, FluxOut* fo) const
  {
    const TaskRec rec = recs[t];
    const BatchJob& J = jobs[job];
    fo->k = rec.k;
    fo->out = rec.out;
    fo->factor = J.v.fp.factor;
    fo->A = J.v.fp.A;
  }
  __device__ void failed(int job) const { atomicAdd(&job_failed[job], 1ull); }
  __device__ const double4* any_ff() const { return jobs[0].ff; }
};

// Contraction job resolver: blk0[j] = first block of job j for this wave count.
struct BatchLumi {
  static constexpr int kWavesPerEu = 4;  // register budget of k_lumi_fp64 (waves per SIMD)
  const BatchJob* __restrict__ jobs;
  const int* __restrict__ blk0;
  int njobs;
  __device__ void get(int blk, int waves, LumiView* out, int* p0) const
  {
    const int j = find_job(blk0, njobs, blk);
    *out = jobs[j].v;
    *p0 = (blk - blk0[j]) * waves;
  }
};

}  // namespace upc

using namespace upc;

struct upc_batch {
  int device = 0;
  int ncu = 256;
  std::vector<upc_ctx*> ctx;
  std::vector<BatchJob> jobs;  // host copy (pointers into the batch buffers)
  long npts = 0, nrows = 0;
  int nparts = 0;
  int lumi_blocks[3] = {0, 0, 0};
  BatchJob* d_jobs = nullptr;
  int* d_pt0 = nullptr;
  int* d_row0 = nullptr;
  int* d_blk0[3] = {nullptr, nullptr, nullptr};
  double* d_pts = nullptr;
  double* d_row_k = nullptr;
  double* d_row_logd = nullptr;
  int* d_row_cnt = nullptr;
  long long* d_row_off = nullptr;
  long long* d_part = nullptr;
  double* d_flux = nullptr;
  long long ntasks = 0;          // QAGS tasks of every job (counted at creation)
  long long work_bytes = 0;      // device memory of the buffers below
  TaskRec* d_tasks = nullptr;
  long long* d_overflow = nullptr;
  unsigned long long* d_counters = nullptr;  // [0] evals [1] bad [2] overflow n [4] tasks [5] cursor
  unsigned long long* d_job_failed = nullptr;  // [njobs] QAGS integrals with ier != 0 per job
  int qags_choice = 2, lumi_choice = 1;  // indices into the candidate lists
  int form = UPC_FORM_FP64;              // contraction form of every job
  int mfma_choice = 2;                   // index into kMfmaTileChoices
  void* mfma[2] = {nullptr, nullptr};   // MFMA states for unpolarised / polarised jobs (largest job's size)
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_q = -1, graph_l = -1, graph_f = -1, graph_m = -1;
  bool outputs_set = false;
};

namespace {

int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

int cu_count(int dev)
{
  int ncu = 0;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  return ncu > 0 ? ncu : 256;
}

void free_batch(upc_batch* b)
{
  mfma_state_free(b->mfma[0]);
  mfma_state_free(b->mfma[1]);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  void* ptrs[] = {b->d_jobs, b->d_pt0, b->d_row0, b->d_blk0[0], b->d_blk0[1], b->d_blk0[2], b->d_pts,
                  b->d_row_k, b->d_row_logd, b->d_row_cnt, b->d_row_off, b->d_part, b->d_flux, b->d_tasks,
                  b->d_overflow, b->d_counters, b->d_job_failed};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (upc_ctx* c : b->ctx) upc_ctx_destroy(c);
  delete b;
}

// The MFMA contraction of every job in turn (the form's kernels per job:
// D from the job's tables, its rows' projection, edge sums, banded GEMM),
// on the flux rows the batch's QAGS launch filled.
int contract_mfma(upc_batch* b, int tile_group, hipStream_t s)
{
  for (int k = 0; k < 2; ++k)
    if (b->mfma[k]) mfma_zero_ksteps(b->mfma[k], s);
  for (size_t j = 0; j < b->jobs.size(); ++j) {
    const BatchJob& J = b->jobs[j];
    if (int rc = mfma_batch_contract(b->ctx[j], b->mfma[J.v.pol ? 1 : 0], J.v, tile_group, s)) return rc;
  }
  return 0;
}

// The whole grid computation of every job (no host synchronisation).
int enqueue(upc_batch* b, hipStream_t s)
{
  const int nj = (int)b->jobs.size();
  hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, s, b->d_counters, 8);
  hipLaunchKernelGGL(k_batch_zero, dim3(nblk(nj, 256)), dim3(256), 0, s, b->d_job_failed, nj);
  hipLaunchKernelGGL(k_batch_points, dim3(nblk(b->npts, 256)), dim3(256), 0, s, b->d_jobs, b->d_pt0, nj,
                     b->npts, b->d_pts);
  hipLaunchKernelGGL(k_batch_rows, dim3(nblk(b->nrows, 256)), dim3(256), 0, s, b->d_jobs, b->d_row0, nj,
                     b->nrows, b->d_row_k, b->d_row_logd, b->d_row_cnt);
  hipLaunchKernelGGL(k_scan_part<>, dim3(b->nparts), dim3(kScanBlock), 0, s, b->d_row_cnt, b->nrows, b->d_part);
  hipLaunchKernelGGL(k_scan_top<>, dim3(1), dim3(kScanBlock), 0, s, b->d_part, b->nparts, b->d_counters + 4);
  hipLaunchKernelGGL(k_scan_apply<>, dim3(b->nparts), dim3(kScanBlock), 0, s, b->d_row_cnt, b->nrows, b->d_part,
                     b->d_row_off);
  hipLaunchKernelGGL(k_batch_point_rows, dim3(nblk(b->nrows * kNB1, 256)), dim3(256), 0, s, b->d_jobs, b->d_row0,
                     nj, b->nrows, b->d_row_k, b->d_row_logd, b->d_row_cnt, b->d_row_off, b->d_flux, b->d_tasks);
  BatchTask ts{b->d_tasks, b->d_jobs, b->d_row0, nj, b->d_job_failed};
  const int qblocks = b->ncu * kQagsBlockChoices[b->qags_choice];
  hipLaunchKernelGGL((k_flux_qags<kQagsCap, BatchTask>), dim3(qblocks), dim3(kQagsBlock), 0, s, ts,
                     b->d_counters + 4, b->d_counters + 5, b->d_flux, (int*)nullptr, b->d_counters, b->d_overflow,
                     kQagsCap);
  hipLaunchKernelGGL((k_flux_qags_wave<BatchTask>), dim3(kQagsWaveBlocks), dim3(64), 0, s, ts, b->d_overflow,
                     b->d_counters, b->d_flux, (int*)nullptr);
  if (b->form != UPC_FORM_FP64) return contract_mfma(b, kMfmaTileChoices[b->mfma_choice], s);
  const int li = b->lumi_choice, w = kLumiWaveChoices[li];
  BatchLumi jr{b->d_jobs, b->d_blk0[li], nj};
  if (w == 2) hipLaunchKernelGGL((k_lumi_fp64<2, BatchLumi>), dim3(b->lumi_blocks[li]), dim3(128), 0, s, jr);
  else if (w == 4) hipLaunchKernelGGL((k_lumi_fp64<4, BatchLumi>), dim3(b->lumi_blocks[li]), dim3(256), 0, s, jr);
  else hipLaunchKernelGGL((k_lumi_fp64<8, BatchLumi>), dim3(b->lumi_blocks[li]), dim3(512), 0, s, jr);
  UPC_HIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int upc_batch_create(const upc_config* cfgs, int njobs, void* stream, upc_batch** out)
{
  if (!cfgs || njobs <= 0 || !out) return set_error("bad arguments"), UPC_EINVAL;
  *out = nullptr;
  hipStream_t s = (hipStream_t)stream;
  auto* b = new upc_batch();
  (void)hipGetDevice(&b->device);
  b->ncu = cu_count(b->device);
  for (int j = 0; j < njobs; ++j) {
    upc_ctx* c = nullptr;
    const int rc = upc_ctx_create(&cfgs[j], (void*)s, &c);
    if (rc) {
      free_batch(b);
      return rc;
    }
    b->ctx.push_back(c);
  }
  // offsets
  std::vector<int> pt0(njobs), row0(njobs), blk0[3];
  long np = 0, nr = 0, blocks[3] = {0, 0, 0};
  for (int j = 0; j < njobs; ++j) {
    const upc_config& g = cfgs[j];
    const long n = (long)g.nm * g.ny;
    const bool mirror = g.ymin == -g.ymax;
    pt0[j] = (int)np;
    row0[j] = (int)nr;
    const int nwork = lumi_work_items((int)n, mirror ? g.ny : 0, mirror);
    for (int i = 0; i < 3; ++i) {
      blk0[i].push_back((int)blocks[i]);
      blocks[i] += nblk(nwork, kLumiWaveChoices[i]);
    }
    np += n;
    nr += mirror ? n + g.nm : 2 * n;
  }
  if (np > INT_MAX || nr > INT_MAX) {
    free_batch(b);
    return set_error("batch too large: more than 2^31 - 1 points or flux rows"), UPC_EINVAL;
  }
  b->npts = np;
  b->nrows = nr;
  b->nparts = nblk(nr, kScanBlock);
  for (int i = 0; i < 3; ++i) b->lumi_blocks[i] = (int)blocks[i];
#define BA(ptr, n)                                                  \
  do {                                                              \
    UPC_HIP(hipMalloc(&(ptr), sizeof(*(ptr)) * (size_t)(n)));       \
    b->work_bytes += (long long)(sizeof(*(ptr)) * (size_t)(n));     \
  } while (0)
  auto alloc = [&]() -> int {
    BA(b->d_jobs, njobs);
    BA(b->d_pt0, njobs);
    BA(b->d_row0, njobs);
    for (int i = 0; i < 3; ++i) BA(b->d_blk0[i], njobs);
    BA(b->d_pts, 4 * np);
    BA(b->d_row_k, nr);
    BA(b->d_row_logd, nr);
    BA(b->d_row_cnt, nr);
    BA(b->d_row_off, nr);
    BA(b->d_part, b->nparts);
    BA(b->d_flux, nr * kNB1);
    BA(b->d_counters, 8);
    BA(b->d_job_failed, njobs);
    return 0;
  };
#undef BA
  if (int rc = alloc()) {
    free_batch(b);
    return rc;
  }
  for (int j = 0; j < njobs; ++j) {
    const upc_ctx* c = b->ctx[j];
    const upc_config& g = cfgs[j];
    const bool mirror = g.ymin == -g.ymax;
    BatchJob J{};
    J.v.pts = b->d_pts + 4L * pt0[j];
    J.v.row_logd = b->d_row_logd + row0[j];
    J.v.flux = b->d_flux + (long)row0[j] * kNB1;
    J.v.fp = flux_params(c, mirror ? g.ny : 0);
    J.v.gaa = c->d_gaa;
    J.v.bk = c->d_bk;
    J.v.sc = c->d_scalars;
    J.v.npts = g.nm * g.ny;
    J.v.mode = g.breakup_mode;
    J.v.pol = g.use_polarized;
    J.v.mirror_out = mirror;
    J.ff = c->d_ff;
    J.nm = g.nm;
    J.ny = g.ny;
    J.mmin = g.mmin;
    J.mmax = g.mmax;
    J.ymin = g.ymin;
    J.ymax = g.ymax;
    J.pt0 = pt0[j];
    J.row0 = row0[j];
    b->jobs.push_back(J);
  }
  auto upload = [&]() -> int {
    UPC_HIP(hipMemcpy(b->d_jobs, b->jobs.data(), sizeof(BatchJob) * njobs, hipMemcpyHostToDevice));
    UPC_HIP(hipMemcpy(b->d_pt0, pt0.data(), sizeof(int) * njobs, hipMemcpyHostToDevice));
    UPC_HIP(hipMemcpy(b->d_row0, row0.data(), sizeof(int) * njobs, hipMemcpyHostToDevice));
    for (int i = 0; i < 3; ++i)
      UPC_HIP(hipMemcpy(b->d_blk0[i], blk0[i].data(), sizeof(int) * njobs, hipMemcpyHostToDevice));
    UPC_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    return 0;
  };
  if (int rc = upload()) {
    free_batch(b);
    return rc;
  }
  // The QAGS task buffers (records, overflow list) are sized by
  // the batch's task count, which the row kernels and the scan give once here.
  auto size_tasks = [&]() -> int {
    const int nj = (int)b->jobs.size();
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, s, b->d_counters, 8);
    hipLaunchKernelGGL(k_batch_points, dim3(nblk(b->npts, 256)), dim3(256), 0, s, b->d_jobs, b->d_pt0, nj,
                       b->npts, b->d_pts);
    hipLaunchKernelGGL(k_batch_rows, dim3(nblk(b->nrows, 256)), dim3(256), 0, s, b->d_jobs, b->d_row0, nj,
                       b->nrows, b->d_row_k, b->d_row_logd, b->d_row_cnt);
    hipLaunchKernelGGL(k_scan_part<>, dim3(b->nparts), dim3(kScanBlock), 0, s, b->d_row_cnt, b->nrows,
                       b->d_part);
    hipLaunchKernelGGL(k_scan_top<>, dim3(1), dim3(kScanBlock), 0, s, b->d_part, b->nparts, b->d_counters + 4);
    UPC_HIP(hipGetLastError());
    unsigned long long total = 0;
    UPC_HIP(hipMemcpyAsync(&total, b->d_counters + 4, sizeof(total), hipMemcpyDeviceToHost, s));
    UPC_HIP(hipStreamSynchronize(s));
    b->ntasks = (long long)total;
    const size_t n = total > 0 ? (size_t)total : 1;
    UPC_HIP(hipMalloc(&b->d_tasks, sizeof(TaskRec) * n));
    UPC_HIP(hipMalloc(&b->d_overflow, sizeof(long long) * n));
    b->work_bytes += (long long)((sizeof(TaskRec) + sizeof(long long)) * n);
    return 0;
  };
  if (int rc = size_tasks()) {
    free_batch(b);
    return rc;
  }
  *out = b;
  return 0;
}

int upc_batch_destroy(upc_batch* b)
{
  if (!b) return set_error("null batch"), UPC_EINVAL;
  (void)hipDeviceSynchronize();
  free_batch(b);
  return 0;
}

int upc_batch_set_outputs(upc_batch* b, double* const* d_L, double* const* d_Ls, double* const* d_Lp)
{
  if (!b) return set_error("null batch"), UPC_EINVAL;
  for (size_t j = 0; j < b->jobs.size(); ++j) {
    BatchJob& J = b->jobs[j];
    if (J.v.pol ? (!d_Ls || !d_Lp || !d_Ls[j] || !d_Lp[j]) : (!d_L || !d_L[j]))
      return set_error("missing output for job " + std::to_string(j)), UPC_EINVAL;
    J.v.L = J.v.pol ? nullptr : d_L[j];
    J.v.Ls = J.v.pol ? d_Ls[j] : nullptr;
    J.v.Lp = J.v.pol ? d_Lp[j] : nullptr;
  }
  // the job table is read by the kernels at run time: captured graphs see the new pointers
  UPC_HIP(hipDeviceSynchronize());
  UPC_HIP(hipMemcpy(b->d_jobs, b->jobs.data(), sizeof(BatchJob) * b->jobs.size(), hipMemcpyHostToDevice));
  b->outputs_set = true;
  b->graph_f = -1;  // the MFMA form's launches take the output pointers as arguments: recapture
  return 0;
}

int upc_batch_set_form(upc_batch* b, int form)
{
  if (!b) return set_error("null batch"), UPC_EINVAL;
  if (form != UPC_FORM_FP64 && form != UPC_FORM_MFMA && form != UPC_FORM_MFMA_FP32)
    return set_error("the batch runs the FP64 or an MFMA form"), UPC_EINVAL;
  if (form != UPC_FORM_FP64) {
    // one state per polarisation (the shared grid's extent differs), sized for its largest job
    size_t rows[2] = {0, 0};
    for (const BatchJob& J : b->jobs) {
      const size_t r = J.v.fp.mirror_ny > 0 ? J.v.npts + J.v.npts / J.v.fp.mirror_ny : 2 * (size_t)J.v.npts;
      rows[J.v.pol ? 1 : 0] = std::max(rows[J.v.pol ? 1 : 0], r);
    }
    for (int k = 0; k < 2; ++k) {
      if (!rows[k]) continue;
      if (!b->mfma[k]) b->mfma[k] = mfma_state_create(0.2);
      if (int rc = mfma_state_ready(b->mfma[k], form, k == 1, rows[k])) return rc;
    }
  }
  b->form = form;
  return 0;
}

int upc_batch_set_mfma_tiles(upc_batch* b, int tile_group)
{
  if (!b) return set_error("null batch"), UPC_EINVAL;
  for (int i = 0; i < 3; ++i)
    if (kMfmaTileChoices[i] == tile_group) {
      b->mfma_choice = i;
      return 0;
    }
  return set_error("unsupported MFMA tile group (1, 2 or 4)"), UPC_EINVAL;
}

int upc_batch_run(upc_batch* b, int use_graph, void* stream)
{
  if (!b) return set_error("null batch"), UPC_EINVAL;
  if (!b->outputs_set) return set_error("upc_batch_set_outputs first"), UPC_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  if (!use_graph) return enqueue(b, s);
  if (!b->exec || b->graph_q != b->qags_choice || b->graph_l != b->lumi_choice || b->graph_f != b->form ||
      b->graph_m != b->mfma_choice) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    UPC_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b, b->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    UPC_HIP(e);
    b->graph = g;
    UPC_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    b->graph_q = b->qags_choice;
    b->graph_l = b->lumi_choice;
    b->graph_f = b->form;
    b->graph_m = b->mfma_choice;
  }
  UPC_HIP(hipGraphLaunch(b->exec, s));
  return 0;
}

int upc_batch_autotune(upc_batch* b, void* stream, upc_tune* h_out)
{
  if (!b) return set_error("null batch"), UPC_EINVAL;
  if (!b->outputs_set) return set_error("upc_batch_set_outputs first"), UPC_ESTATE;
  hipStream_t s = (hipStream_t)stream;
  upc_tune t{};
  int rc = enqueue(b, s);  // flux rows, tasks and a warm run
  if (rc) return rc;
  hipEvent_t e0, e1;
  UPC_HIP(hipEventCreate(&e0));
  UPC_HIP(hipEventCreate(&e1));
  const int nj = (int)b->jobs.size();
  BatchTask ts{b->d_tasks, b->d_jobs, b->d_row0, nj, b->d_job_failed};
  for (int i = 0; i < 3; ++i) {
    // QAGS alone: clear the counters except the task total, then time
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, s, b->d_counters, 4);
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(64), 0, s, b->d_counters + 5, 1);
    UPC_HIP(hipEventRecord(e0, s));
    const int qblocks = b->ncu * kQagsBlockChoices[i];
    hipLaunchKernelGGL((k_flux_qags<kQagsCap, BatchTask>), dim3(qblocks), dim3(kQagsBlock), 0, s, ts,
                       b->d_counters + 4, b->d_counters + 5, b->d_flux, (int*)nullptr, b->d_counters, b->d_overflow,
                       kQagsCap);
    UPC_HIP(hipEventRecord(e1, s));
    UPC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    UPC_HIP(hipEventElapsedTime(&ms, e0, e1));
    t.qags_candidates[i] = kQagsBlockChoices[i];
    t.qags_ms[i] = ms;
  }
  // restore the overflow results (the timed passes re-ran every task)
  hipLaunchKernelGGL((k_flux_qags_wave<BatchTask>), dim3(kQagsWaveBlocks), dim3(64), 0, s, ts, b->d_overflow,
                     b->d_counters, b->d_flux, (int*)nullptr);
  for (int i = 0; i < 3 && b->form != UPC_FORM_FP64; ++i) {
    // the MFMA form's tile shape: column tiles per A-fragment load (results identical)
    UPC_HIP(hipEventRecord(e0, s));
    rc = contract_mfma(b, kMfmaTileChoices[i], s);
    if (rc) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return rc;
    }
    UPC_HIP(hipEventRecord(e1, s));
    UPC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    UPC_HIP(hipEventElapsedTime(&ms, e0, e1));
    t.mfma_candidates[i] = kMfmaTileChoices[i];
    t.mfma_ms[i] = ms;
  }
  for (int i = 0; i < 3 && b->form == UPC_FORM_FP64; ++i) {
    BatchLumi jr{b->d_jobs, b->d_blk0[i], nj};
    UPC_HIP(hipEventRecord(e0, s));
    const int w = kLumiWaveChoices[i];
    if (w == 2) hipLaunchKernelGGL((k_lumi_fp64<2, BatchLumi>), dim3(b->lumi_blocks[i]), dim3(128), 0, s, jr);
    else if (w == 4) hipLaunchKernelGGL((k_lumi_fp64<4, BatchLumi>), dim3(b->lumi_blocks[i]), dim3(256), 0, s, jr);
    else hipLaunchKernelGGL((k_lumi_fp64<8, BatchLumi>), dim3(b->lumi_blocks[i]), dim3(512), 0, s, jr);
    UPC_HIP(hipEventRecord(e1, s));
    UPC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    UPC_HIP(hipEventElapsedTime(&ms, e0, e1));
    t.lumi_candidates[i] = w;
    t.lumi_ms[i] = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->qags_choice = (int)(std::min_element(t.qags_ms, t.qags_ms + 3) - t.qags_ms);
  if (b->form == UPC_FORM_FP64) b->lumi_choice = (int)(std::min_element(t.lumi_ms, t.lumi_ms + 3) - t.lumi_ms);
  else b->mfma_choice = (int)(std::min_element(t.mfma_ms, t.mfma_ms + 3) - t.mfma_ms);
  t.qags_blocks_per_cu = kQagsBlockChoices[b->qags_choice];
  t.lumi_waves = kLumiWaveChoices[b->lumi_choice];
  t.mfma_tile_group = kMfmaTileChoices[b->mfma_choice];
  t.form = b->form;
  if (h_out) *h_out = t;
  return 0;
}

int upc_batch_set_tune(upc_batch* b, int qags_blocks_per_cu, int lumi_waves)
{
  if (!b) return set_error("null batch"), UPC_EINVAL;
  int qi = -1, li = -1;
  for (int i = 0; i < 3; ++i) {
    if (kQagsBlockChoices[i] == qags_blocks_per_cu) qi = i;
    if (kLumiWaveChoices[i] == lumi_waves) li = i;
  }
  if (qi < 0 || li < 0) return set_error("unsupported launch shape"), UPC_EINVAL;
  b->qags_choice = qi;
  b->lumi_choice = li;
  return 0;
}

int upc_batch_job_failures(upc_batch* b, void* stream, long long* h_failed)
{
  if (!b || !h_failed) return set_error("null pointer"), UPC_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  UPC_HIP(hipMemcpyAsync(h_failed, b->d_job_failed, sizeof(long long) * b->jobs.size(), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipStreamSynchronize(s));
  return 0;
}

int upc_batch_info(upc_batch* b, void* stream, upc_batch_stats* h_out)
{
  if (!b || !h_out) return set_error("null pointer"), UPC_EINVAL;
  unsigned long long h[8];
  hipStream_t s = (hipStream_t)stream;
  UPC_HIP(hipMemcpyAsync(h, b->d_counters, sizeof(h), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipStreamSynchronize(s));
  h_out->njobs = (int)b->jobs.size();
  h_out->points = b->npts;
  h_out->rows = b->nrows;
  h_out->qags_calls = (long long)h[4];
  h_out->qags_evals = (long long)h[0];
  h_out->qags_bad = (long long)h[1];
  h_out->qags_blocks_per_cu = kQagsBlockChoices[b->qags_choice];
  h_out->lumi_waves = kLumiWaveChoices[b->lumi_choice];
  h_out->tasks = b->ntasks;
  h_out->work_bytes = b->work_bytes;
  h_out->form = b->form;
  h_out->mfma_tile_group = kMfmaTileChoices[b->mfma_choice];
  h_out->mfma_fp64_items = 0;
  if (b->form != UPC_FORM_FP64)
    for (int k = 0; k < 2; ++k) {
      if (!b->mfma[k]) continue;
      unsigned long long c[2];
      if (int rc = mfma_read_counts(b->mfma[k], c, s)) return rc;
      h_out->mfma_fp64_items += (long long)c[1];
    }
  return 0;
}

}  // extern "C"
