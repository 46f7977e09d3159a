// C ABI of libupcgen_amd (declared in include/upc_lumi.h).
#include "upc_internal.h"

#include <cmath>
#include <cstring>
#include <new>
#include <string>

namespace upc {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char* what)
{
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? UPC_ENOMEM : UPC_EDEVICE;
}

PhaseScope::PhaseScope(upc_ctx* c_, hipStream_t s_, int phase_) : c(c_), s(s_), phase(phase_)
{
  if (!c || !c->timing) return;
  hipEvent_t b;
  if (hipEventCreate(&b) != hipSuccess) return;
  if (hipEventCreate(&end) != hipSuccess) {
    (void)hipEventDestroy(b);
    end = nullptr;
    return;
  }
  (void)hipEventRecord(b, s);
  c->recs.push_back(TimedRec{phase, b, end});
}

PhaseScope::~PhaseScope()
{
  if (end) (void)hipEventRecord(end, s);
}

// MFMA / surrogate forms live in their own translation units.
int lumi_grid_mfma(upc_ctx* c, int prec, double* d_L, double* d_Ls, double* d_Lp, hipStream_t s);
int mfma_set_spacing(upc_ctx* c, double h);
unsigned long long mfma_last_ksteps(upc_ctx* c);
unsigned long long mfma_last_fp64_items(upc_ctx* c);
int mfma_set_cancel_limit(upc_ctx* c, double limit);
double mfma_spacing(upc_ctx* c);
void surrogate_counts(upc_ctx* c, long long* served, long long* fallback);
int lumi_grid_surrogate(upc_ctx* c, double* d_L, double* d_Ls, double* d_Lp, hipStream_t s);
void free_mfma(upc_ctx* c);
void free_surrogate(upc_ctx* c);

static int validate(const upc_config* g)
{
  if (!g) return set_error("null config"), UPC_EINVAL;
  if (g->Z <= 0 || g->A <= 0 || g->Z > g->A) return set_error("need 0 < Z <= A"), UPC_EINVAL;
  if (!(g->R > 0) || !(g->a > 0)) return set_error("Woods-Saxon R and a must be > 0"), UPC_EINVAL;
  if (!(g->sqrts > 0)) return set_error("SQRTS must be > 0"), UPC_EINVAL;
  if (g->breakup_mode < 1 || g->breakup_mode > 4) return set_error("BREAKUP_MODE must be 1..4"), UPC_EINVAL;
  if (g->nm <= 0 || g->ny <= 0 || g->nz <= 0) return set_error("bin counts must be > 0"), UPC_EINVAL;
  if (!(g->mmax > g->mmin) || !(g->ymax > g->ymin) || !(g->zmax > g->zmin))
    return set_error("grid ranges must be increasing"), UPC_EINVAL;
  if (!(g->mmin > 0)) return set_error("MMIN must be > 0"), UPC_EINVAL;
  if ((long long)g->nm * g->ny > 16000000LL) return set_error("nm*ny above 1.6e7 points"), UPC_EINVAL;
  if (g->proc_id != 11 && g->proc_id != 13 && g->proc_id != 15)
    return set_error("PROC_ID must be 11, 13 or 15 (dilepton processes)"), UPC_EINVAL;
  return 0;
}

static int alloc_host_copies(upc_ctx* c)
{
  UPC_HIP(hipHostMalloc((void**)&c->hp_scalars, sizeof(DevScalars), hipHostMallocDefault));
  UPC_HIP(hipHostMalloc((void**)&c->hp_stats, sizeof(unsigned long long) * 8, hipHostMallocDefault));
  UPC_HIP(hipEventCreateWithFlags(&c->ev_scalars, hipEventDisableTiming));
  UPC_HIP(hipEventCreateWithFlags(&c->ev_stats, hipEventDisableTiming));
  return 0;
}

static void free_host_copies(upc_ctx* c)
{
  for (hipEvent_t e : {c->ev_scalars, c->ev_stats})
    if (e) {
      (void)hipEventSynchronize(e);
      (void)hipEventDestroy(e);
    }
  if (c->hp_scalars) (void)hipHostFree(c->hp_scalars);
  if (c->hp_stats) (void)hipHostFree(c->hp_stats);
  c->ev_scalars = c->ev_stats = nullptr;
  c->hp_scalars = nullptr;
  c->hp_stats = nullptr;
}

int host_copies(upc_ctx* c)
{
  if (c->scalars_pending) {
    UPC_HIP(hipEventSynchronize(c->ev_scalars));
    c->h_scalars = *c->hp_scalars;
    c->scalars_pending = false;
  }
  if (c->stats_pending) {
    UPC_HIP(hipEventSynchronize(c->ev_stats));
    c->last_qags_evals = (long long)c->hp_stats[0];
    c->last_qags_bad = (long long)c->hp_stats[1];
    c->last_qags_calls = (long long)c->hp_stats[4];
    c->stats_pending = false;
  }
  return 0;
}

}  // namespace upc

using namespace upc;

static inline hipStream_t S(void* s) { return (hipStream_t)s; }

extern "C" {

const char* upc_last_error(void) { return g_last_error.c_str(); }

int upc_device_count(int* h_count)
{
  if (!h_count) return set_error("null pointer"), UPC_EINVAL;
  int n = 0;
  UPC_HIP(hipGetDeviceCount(&n));
  *h_count = n;
  return 0;
}

int upc_ctx_create(const upc_config* cfg, void* stream, upc_ctx** out)
{
  if (!out) return set_error("null output pointer"), UPC_EINVAL;
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  int ndev = 0;
  UPC_HIP(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return set_error("no HIP device"), UPC_EDEVICE;
  upc_ctx* c = new (std::nothrow) upc_ctx();
  if (!c) return set_error("host allocation failed"), UPC_ENOMEM;
  c->cfg = *cfg;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return set_error("hipGetDevice failed"), UPC_EDEVICE;
  }
  // src/UpcGenerator.cpp:188-192 and src/UpcCrossSection.cpp:120-121
  c->g1 = cfg->sqrts / (2. * kMProt);
  c->g2 = cfg->sqrts / (2. * kMProt);
  c->factor = cfg->Z * cfg->Z * kAlpha / M_PI / M_PI / kHc / kHc;
  c->mNucl = (cfg->Z * kMProt + (cfg->A - cfg->Z) * kMNeut) / cfg->A;
  rc = alloc_tables(c);
  if (!rc) rc = alloc_host_copies(c);
  if (!rc) rc = build_tables(c, S(stream));
  if (rc) {
    upc_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int upc_ctx_rebuild_tables(upc_ctx* ctx, void* stream)
{
  if (!ctx) return set_error("null context"), UPC_EINVAL;
  return build_tables(ctx, S(stream));
}

int upc_ctx_destroy(upc_ctx* ctx)
{
  if (!ctx) return 0;
  for (TimedRec& r : ctx->recs) {
    (void)hipEventSynchronize(r.end);
    (void)hipEventDestroy(r.beg);
    (void)hipEventDestroy(r.end);
  }
  free_host_copies(ctx);
  free_work(ctx);
  free_tables(ctx);
  free_mfma(ctx);
  free_surrogate(ctx);
  free_events(ctx);
  delete ctx;
  return 0;
}

int upc_ctx_info(upc_ctx* c, upc_info* h)
{
  if (!c || !h) return set_error("null pointer"), UPC_EINVAL;
  if (int rc = host_copies(c)) return rc;
  h->g1 = c->g1;
  h->g2 = c->g2;
  h->factor = c->factor;
  h->mNucl = c->mNucl;
  h->rho0 = c->h_scalars.rho0;
  h->csNN = c->h_scalars.csNN;
  h->p20 = c->h_scalars.p20;
  h->bskip = c->h_scalars.bskip;
  h->qags_calls = c->last_qags_calls;
  h->qags_bad = c->last_qags_bad;
  h->qags_evals = c->last_qags_evals;
  h->mfma_ksteps = (long long)mfma_last_ksteps(c);
  h->mfma_fp64_items = (long long)mfma_last_fp64_items(c);
  h->mfma_spacing = mfma_spacing(c);
  surrogate_counts(c, &h->surrogate_points, &h->surrogate_fallback);
  return 0;
}

int upc_ctx_set_mfma_spacing(upc_ctx* c, double h_fm)
{
  if (!c) return set_error("null context"), UPC_EINVAL;
  return mfma_set_spacing(c, h_fm);
}

int upc_ctx_set_mfma_cancel_limit(upc_ctx* c, double limit)
{
  if (!c) return set_error("null context"), UPC_EINVAL;
  return mfma_set_cancel_limit(c, limit);
}

int upc_ctx_set_qags_cap(upc_ctx* c, int cap)
{
  if (!c) return set_error("null context"), UPC_EINVAL;
  if (cap < 1 || cap > kQagsCap) return set_error("QAGS fast-path cap must be in [1, " + std::to_string(kQagsCap) + "]"), UPC_EINVAL;
  c->qags_cap = cap;
  return 0;
}

int upc_ctx_set_timing(upc_ctx* c, int enable)
{
  if (!c) return set_error("null context"), UPC_EINVAL;
  c->timing = enable != 0;
  return 0;
}

int upc_ctx_timings(upc_ctx* c, upc_timings* h, int reset)
{
  if (!c || !h) return set_error("null pointer"), UPC_EINVAL;
  for (TimedRec& r : c->recs) {
    float ms = 0.f;
    UPC_HIP(hipEventSynchronize(r.end));
    UPC_HIP(hipEventElapsedTime(&ms, r.beg, r.end));
    c->t_ms[r.phase] += ms;
    c->t_n[r.phase] += 1;
    (void)hipEventDestroy(r.beg);
    (void)hipEventDestroy(r.end);
  }
  c->recs.clear();
  for (int i = 0; i < UPC_NPHASES; ++i) {
    h->ms[i] = c->t_ms[i];
    h->launches[i] = c->t_n[i];
  }
  if (reset) {
    for (int i = 0; i < UPC_NPHASES; ++i) {
      c->t_ms[i] = 0;
      c->t_n[i] = 0;
    }
  }
  return 0;
}

int upc_flux_point(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out, void* stream)
{
  if (!c || n < 0 || (n && (!d_b || !d_k || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_flux_point(c, d_b, d_k, n, d_out, S(stream));
}

int upc_flux_form(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out, int* d_ier,
                  void* stream)
{
  if (!c || n < 0 || (n && (!d_b || !d_k || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_flux_form(c, d_b, d_k, n, d_out, d_ier, S(stream));
}

int upc_bessel_j1(const double* d_x, int n, double* d_out, void* stream)
{
  if (n < 0 || (n && (!d_x || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_bessel_j1(d_x, n, d_out, S(stream));
}

int upc_form_factor(upc_ctx* c, const double* d_q2, int n, int tabulated, double* d_out, void* stream)
{
  if (!c || n < 0 || (n && (!d_q2 || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_form_factor(c, d_q2, n, tabulated, d_out, S(stream));
}

int upc_breakup_prob(upc_ctx* c, const double* d_b, int n, int mode, int tabulated, double* d_out, void* stream)
{
  if (!c || n < 0 || (n && (!d_b || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_breakup(c, d_b, n, mode, tabulated, d_out, S(stream));
}

int upc_nuclear_overlap(upc_ctx* c, const double* d_b, int n, int which, double* d_out, void* stream)
{
  if (!c || n < 0 || (n && (!d_b || !d_out)) || (which != 0 && which != 1))
    return set_error("bad arguments"), UPC_EINVAL;
  return launch_overlap(c, d_b, n, which, d_out, S(stream));
}

int upc_lumi_points(upc_ctx* c, const double* d_M, const double* d_Y, int n, double* d_L, double* d_Ls,
                    double* d_Lp, void* stream)
{
  if (!c || n < 0) return set_error("bad arguments"), UPC_EINVAL;
  if (n == 0) return 0;
  if (!d_M || !d_Y) return set_error("null M/Y"), UPC_EINVAL;
  if (c->cfg.use_polarized ? (!d_Ls || !d_Lp) : !d_L) return set_error("missing output"), UPC_EINVAL;
  // d_M and d_Y need not be contiguous with each other: stage them.
  double* tmp = nullptr;
  UPC_HIP(hipMallocAsync((void**)&tmp, sizeof(double) * 2 * n, S(stream)));
  UPC_HIP(hipMemcpyAsync(tmp, d_M, sizeof(double) * n, hipMemcpyDeviceToDevice, S(stream)));
  UPC_HIP(hipMemcpyAsync(tmp + n, d_Y, sizeof(double) * n, hipMemcpyDeviceToDevice, S(stream)));
  const int rc = lumi_points_fp64(c, tmp, n, d_L, d_Ls, d_Lp, S(stream));
  (void)hipFreeAsync(tmp, S(stream));
  return rc;
}

int upc_lumi_grid(upc_ctx* c, int form, double* d_L, double* d_Ls, double* d_Lp, void* stream)
{
  if (!c) return set_error("null context"), UPC_EINVAL;
  if (c->cfg.use_polarized ? (!d_Ls || !d_Lp) : !d_L) return set_error("missing output"), UPC_EINVAL;
  switch (form) {
    case UPC_FORM_FP64: {
      int rc = make_grid_points(c, S(stream));
      if (rc) return rc;
      return lumi_points_fp64(c, nullptr, c->cfg.nm * c->cfg.ny, d_L, d_Ls, d_Lp, S(stream));
    }
    case UPC_FORM_MFMA:
    case UPC_FORM_MFMA_FP32:
      return lumi_grid_mfma(c, form, d_L, d_Ls, d_Lp, S(stream));
    case UPC_FORM_SURROGATE:
      return lumi_grid_surrogate(c, d_L, d_Ls, d_Lp, S(stream));
    default:
      return set_error("unknown form"), UPC_EINVAL;
  }
}

int upc_photon_flux(upc_ctx* c, const double* d_M, const double* d_Y, int n, double* d_out, void* stream)
{
  if (!c || n < 0 || (n > 0 && (!d_M || !d_Y || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_photon_flux(c, d_M, d_Y, n, d_out, S(stream));
}

int upc_nuc_cs_y(upc_ctx* c, double mpart, const double* d_cs_pos, const double* d_cs_neg, double* d_cs,
                 double* d_ratio, double* h_tot, void* stream)
{
  if (!c || !d_cs_pos || !d_cs_neg || !d_cs || !d_ratio || !h_tot || !(mpart > 0))
    return set_error("bad arguments"), UPC_EINVAL;
  return launch_nuc_cs_y(c, mpart, d_cs_pos, d_cs_neg, d_cs, d_ratio, h_tot, S(stream));
}

int upc_cs_zm(upc_ctx* c, double zmin, double zmax, int nz, double mmin, double mmax, int nm, int flag,
              double* d_zm, void* stream)
{
  if (!c || !d_zm || flag < 0 || flag > 2 || nz <= 0 || nm <= 0 || !(zmax > zmin) || !(mmax > mmin))
    return set_error("bad arguments"), UPC_EINVAL;
  return launch_cs_zm(c, zmin, zmax, nz, mmin, mmax, nm, flag, d_zm, S(stream));
}

int upc_cs_m(upc_ctx* c, int flag, double* d_cs, void* stream)
{
  if (!c || !d_cs || flag < 0 || flag > 2) return set_error("bad arguments"), UPC_EINVAL;
  return launch_cs_m(c, flag, d_cs, S(stream));
}

int upc_elem_cs(upc_ctx* c, int flag, const double* d_m, const double* d_z, int n, double* d_out, void* stream)
{
  if (!c || n < 0 || flag < 0 || flag > 5 || (n && (!d_m || !d_out || (flag >= 3 && !d_z))))
    return set_error("bad arguments"), UPC_EINVAL;
  return launch_elem_cs(c, flag, d_m, d_z, n, d_out, S(stream));
}

int upc_nuc_cs_ym(upc_ctx* c, const double* d_L, const double* d_Ls, const double* d_Lp, double* d_cs,
                  double* d_ratio, double* h_tot_cs, void* stream)
{
  if (!c || !d_cs) return set_error("bad arguments"), UPC_EINVAL;
  if (c->cfg.use_polarized ? (!d_Ls || !d_Lp || !d_ratio) : !d_L) return set_error("missing input"), UPC_EINVAL;
  return launch_nuc_cs_ym(c, d_L, d_Ls, d_Lp, d_cs, d_ratio, h_tot_cs, S(stream));
}

int upc_pdf_init(const double* d_bins, int nrows, int n, double* d_sum, void* stream)
{
  if (nrows < 0 || n < 0 || ((nrows && n) && (!d_bins || !d_sum))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_pdf_init(d_bins, nrows, n, d_sum, S(stream));
}

int upc_sample_events(upc_ctx* c, const double* d_sum_ym, const double* d_sum_z, const double* d_sum_zs,
                      const double* d_sum_zp, const double* d_ratio, unsigned long long seed,
                      unsigned long long offset, long long n, double* d_y, double* d_m, double* d_z,
                      void* stream)
{
  if (!c || n < 0 || (n && (!d_sum_ym || !d_y || !d_m || !d_z))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_sample_events(c, d_sum_ym, d_sum_z, d_sum_zs, d_sum_zp, d_ratio, seed, offset, n, d_y, d_m, d_z,
                              S(stream));
}

int upc_photon_pt_tables(upc_ctx* c, double gtot, void* stream)
{
  if (!c || !(gtot >= 1.0)) return set_error("bad arguments"), UPC_EINVAL;
  return build_pt_tables(c, gtot, S(stream));
}

int upc_photon_pt(upc_ctx* c, const double* d_E, long long n, unsigned long long seed, double* d_out, void* stream)
{
  if (!c || n < 0 || (n && (!d_E || !d_out))) return set_error("bad arguments"), UPC_EINVAL;
  return launch_photon_pt(c, d_E, n, seed, d_out, S(stream));
}

int upc_generate_pairs(upc_ctx* c, const double* d_sum_ym, const double* d_sum_z, const double* d_sum_zs,
                       const double* d_sum_zp, const double* d_ratio, const upc_event_opts* opts,
                       unsigned long long seed, unsigned long long offset, long long n, double* d_p, int* d_pdg,
                       int* d_pass, void* stream)
{
  if (!c || !opts || n < 0 || (n && (!d_sum_ym || !d_p || !d_pdg || !d_pass)))
    return set_error("bad arguments"), UPC_EINVAL;
  return launch_generate_pairs(c, d_sum_ym, d_sum_z, d_sum_zs, d_sum_zp, d_ratio, opts, seed, offset, n, d_p, d_pdg,
                               d_pass, S(stream));
}

}  // extern "C"
