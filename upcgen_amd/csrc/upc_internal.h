// Internal declarations shared by the HIP translation units of libupcgen_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/upc_lumi.h"
#include "upc_device.h"

namespace upc {

// Device-resident scalars computed by the table kernels.
struct DevScalars {
  double rho0;
  double csNN;
  double p20;       // breakup spline at b = 20 (1 for mode 1)
  double bskip;     // |G_AA spline| < p20 2^-71 on [0, bskip)
  double zcon;      // calcBreakupProb statics
  double o0;
  int nee;          // index of the sentinel energy
  int pad;
};

// Fast per-lane QAGS interval list (the flux integrals of the 1001x121
// Pb-Pb grid stop by 15 intervals); integrals that need more continue in the
// 1000-interval kernel.  UPC_QAGS_FAST_CAP (2..16) sizes it at build time.
#ifndef UPC_QAGS_FAST_CAP
#define UPC_QAGS_FAST_CAP 16
#endif
constexpr int kQagsCap = UPC_QAGS_FAST_CAP;
constexpr int kMaxEE = 1024;  // calcBreakupProb energy table capacity (reference: 10001)

// Breakup knots actually needed: the luminosity reads P(min(b, 20)); the
// natural-spline coupling decays by 2-sqrt(3) per knot, so knots past
// index 20000 + kBkGuard change the b <= 20 interpolant by < 1e-35.
constexpr int kBkGuard = 96;
constexpr int kNBkUsed = 20000 + kBkGuard;

struct LumiWork {
  size_t rows_cap = 0;
  double* d_row_k = nullptr;      // photon energy of each flux row
  double* d_row_logd = nullptr;   // log step of the row's b grid
  int* d_row_cnt = nullptr;       // number of QAGS nodes in the row
  long long* d_row_off = nullptr; // exclusive scan of d_row_cnt
  double* d_flux = nullptr;       // [rows][120]
  size_t tasks_cap = 0;
  void* d_tasks = nullptr;        // QAGS task records (b, k, flux slot); per-point calls: point indices
  unsigned long long* d_counters = nullptr;  // [0] task cursor, [1] qags bad, [2] evals, [3] overflow n
  long long* d_overflow = nullptr;  // tasks that exceeded the fast workspace
  double* d_points = nullptr;     // (M, Y, dm, dy) per point
  long long* d_part = nullptr;    // per-block partial sums of the row-count scan
  size_t points_cap = 0;
};

// Buffers of one per-point form-factor flux launch (upc_flux_form).
struct FluxFormBufs {
  int* idx;                      // [n] compacted indices of the points that need QAGS
  long long* overflow;           // [n] tasks that outgrew the fast interval list
  unsigned long long* counters;  // [8] evaluations, failed, overflow count, -, tasks, cursor
};

// One timed phase: the event pair recorded around it when timing is enabled.
struct TimedRec {
  int phase;
  hipEvent_t beg, end;
};

}  // namespace upc

struct upc_ctx {
  upc_config cfg;
  bool timing = false;
  std::vector<upc::TimedRec> recs;
  double t_ms[UPC_NPHASES] = {};
  long long t_n[UPC_NPHASES] = {};
  int device = 0;
  double g1 = 0, g2 = 0, factor = 0, mNucl = 0;
  upc::DevScalars h_scalars{};      // host copy after init (see host_copies())
  upc::DevScalars* d_scalars = nullptr;
  // Device-to-host copies that only upc_ctx_info reads are queued into pinned
  // memory behind an event instead of synchronising the stream; host_copies()
  // waits for them when info is asked for.
  upc::DevScalars* hp_scalars = nullptr;
  unsigned long long* hp_stats = nullptr;  // the QAGS counter block of the last flux launch
  hipEvent_t ev_scalars = nullptr, ev_stats = nullptr;
  bool scalars_pending = false, stats_pending = false;
  double4* d_ta = nullptr;          // 199 intervals
  double4* d_gaa = nullptr;         // 199 intervals
  double* d_ff_y = nullptr;         // 10^6 knot values
  double* d_ff_c = nullptr;         // 10^6 second-derivative coefficients
  double4* d_ff = nullptr;          // 10^6 - 1 intervals
  double* d_bk_y = nullptr;         // kNBkUsed knot values
  double* d_bk_c = nullptr;
  double4* d_bk = nullptr;          // kNBkUsed - 1 intervals
  double* d_ee = nullptr;           // calcBreakupProb tables ee[], se[]
  double* d_se = nullptr;
  double* d_spline_scratch = nullptr;  // windowed tridiagonal solver scratch (form-factor spline)
  size_t spline_scratch_elems = 0;
  double* d_bk_scratch = nullptr;      // the same for the breakup spline, built concurrently
  size_t bk_scratch_elems = 0;
  double* d_gaa_scratch = nullptr;     // G_AA knot values
  // The three table chains (G_AA, form factor, breakup) are independent: two
  // side streams (default priority) run two of them beside the caller's.
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  upc::LumiWork work;
  long long last_qags_calls = 0, last_qags_bad = 0, last_qags_evals = 0;
  int qags_cap = upc::kQagsCap;  // intervals of the fast QAGS path before the 1000-interval kernel
  // MFMA / surrogate state is attached by their translation units.
  void* mfma = nullptr;
  void* surrogate = nullptr;
  void* events = nullptr;
};

namespace upc {

void set_error(const std::string& msg);
int host_copies(upc_ctx* c);  // completes the queued device-to-host copies

// Records an event pair around a phase on `s` when timing is enabled.
struct PhaseScope {
  upc_ctx* c;
  hipStream_t s;
  int phase;
  hipEvent_t end = nullptr;
  PhaseScope(upc_ctx* c_, hipStream_t s_, int phase_);
  ~PhaseScope();
};
int hip_fail(hipError_t e, const char* what);
#define UPC_HIP(call)                                          \
  do {                                                         \
    hipError_t e_ = (call);                                    \
    if (e_ != hipSuccess) return ::upc::hip_fail(e_, #call);   \
  } while (0)

// upc_tables.hip
int build_tables(upc_ctx* c, hipStream_t s);
int alloc_tables(upc_ctx* c);
void free_tables(upc_ctx* c);
int launch_form_factor(upc_ctx* c, const double* d_q2, int n, int tabulated, double* d_out,
                       hipStream_t s);
int launch_breakup(upc_ctx* c, const double* d_b, int n, int mode, int tabulated, double* d_out,
                   hipStream_t s);
int launch_overlap(upc_ctx* c, const double* d_b, int n, int which, double* d_out, hipStream_t s);

// upc_flux_lumi.hip
int lumi_points_fp64(upc_ctx* c, const double* d_pts, int npts, double* d_L, double* d_Ls,
                     double* d_Lp, hipStream_t s);
int make_grid_points(upc_ctx* c, hipStream_t s);
// The FP64 grid of the bins with d_need_bin[i] != 0 only (flux rows and
// contraction items they need; other bins of d_Lt are undefined).
int lumi_grid_subset_fp64(upc_ctx* c, const unsigned char* d_need_bin, double* d_Lt, hipStream_t s);
int launch_flux_point(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out,
                      hipStream_t s);
int launch_flux_form(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out,
                     int* d_ier, hipStream_t s);
int enqueue_flux_form(upc_ctx* c, const double* d_b, const double* d_k, int n, double* d_out, int* d_ier,
                      const FluxFormBufs& w, hipStream_t s);
void free_work(upc_ctx* c);
int launch_bessel_j1(const double* d_x, int n, double* d_out, hipStream_t s);
int launch_photon_flux(upc_ctx* c, const double* d_M, const double* d_Y, int n, double* d_out, hipStream_t s);
int launch_nuc_cs_y(upc_ctx* c, double mpart, const double* d_cs_pos, const double* d_cs_neg, double* d_cs,
                    double* d_ratio, double* h_tot, hipStream_t s);

// upc_xsec.hip
int launch_cs_zm(upc_ctx* c, double zmin, double zmax, int nz, double mmin, double mmax, int nm, int flag,
                 double* d_zm, hipStream_t s);
int launch_cs_m(upc_ctx* c, int flag, double* d_cs, hipStream_t s);
int launch_elem_cs(upc_ctx* c, int flag, const double* d_m, const double* d_z, int n, double* d_out,
                   hipStream_t s);
int launch_nuc_cs_ym(upc_ctx* c, const double* d_L, const double* d_Ls, const double* d_Lp,
                     double* d_cs, double* d_ratio, double* h_tot, hipStream_t s);
int launch_pdf_init(const double* d_bins, int nrows, int n, double* d_sum, hipStream_t s);
int launch_sample_events(upc_ctx* c, const double* d_sum_ym, const double* d_sum_z,
                         const double* d_sum_zs, const double* d_sum_zp, const double* d_ratio,
                         unsigned long long seed, unsigned long long offset, long long n,
                         double* d_y, double* d_m, double* d_z, hipStream_t s);

// upc_events.hip
int build_pt_tables(upc_ctx* c, double gtot, hipStream_t s);
void free_events(upc_ctx* c);
int launch_photon_pt(upc_ctx* c, const double* d_E, long long n, unsigned long long seed, double* d_out,
                     hipStream_t s);
int launch_generate_pairs(upc_ctx* c, const double* d_sum_ym, const double* d_sum_z, const double* d_sum_zs,
                          const double* d_sum_zp, const double* d_ratio, const upc_event_opts* opts,
                          unsigned long long seed, unsigned long long offset, long long n, double* d_p, int* d_pdg,
                          int* d_pass, hipStream_t s);

}  // namespace upc
