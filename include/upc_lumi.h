/*
 * upc_lumi.h -- C ABI of the MI355X-native two-photon luminosity and nuclear
 * cross-section engine (libupcgen_amd.so).
 *
 * Drop-in boundary for upcgen's UpcCrossSection / UpcGenerator path.  The
 * reference exposes this path only as a C++ class API (include/UpcCrossSection.h
 * :174-239, include/UpcGenerator.h:75-128); every entry point below names the
 * member function it replaces.  The rebuilt C++ facade (upcgen_amd/csrc/
 * UpcCrossSection.h, UpcGenerator.h) is a thin layer over these calls.
 *
 * Conventions
 *   - plain C types only; `stream` is a hipStream_t passed as void* (NULL =
 *     the null stream);
 *   - pointers named d_* are device (HBM) pointers owned by the caller, h_*
 *     are host pointers;
 *   - every function returns 0 on success or a negative UPC_E* code; the
 *     message of the last failure on the calling thread is upc_last_error();
 *   - a context is one configuration (one reference process); contexts are
 *     independent and may live on the same device;
 *   - calls that return host values (h_*) synchronise `stream`.
 */
#ifndef UPC_LUMI_H
#define UPC_LUMI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UPC_OK 0
#define UPC_EINVAL (-1)   /* bad argument or configuration */
#define UPC_EDEVICE (-2)  /* HIP runtime error */
#define UPC_ENOMEM (-3)   /* device allocation failed */
#define UPC_ESTATE (-4)   /* call not valid in the context's state */

/* Luminosity compute forms. */
#define UPC_FORM_FP64 0       /* reference quadrature, FP64 on the GPU */
#define UPC_FORM_MFMA 1       /* FP64 flux rows, shared-grid W.D.W^T on MFMA (split-bf16, FP32 acc) */
#define UPC_FORM_SURROGATE 2  /* MLP surrogate with quadrature fallback */
#define UPC_FORM_MFMA_FP32 3  /* as UPC_FORM_MFMA with FP32 MFMA operands */

/* Configuration: the parameters.in keys that reach the luminosity path
 * (src/UpcGenerator.cpp:183-305; defaults of include/UpcCrossSection.h:72-164). */
typedef struct upc_config {
  int Z, A;              /* NUCLEUS_Z, NUCLEUS_A */
  double R, a;           /* WS_R, WS_A [fm] */
  double sqrts;          /* SQRTS [GeV]; g1 = g2 = sqrts / (2 m_p) */
  int flux_point;        /* FLUX_POINT */
  int breakup_mode;      /* BREAKUP_MODE 1..4 */
  int use_polarized;     /* USE_POLARIZED_CS */
  int nm; double mmin, mmax;   /* BINS_M, MMIN, MMAX */
  int ny; double ymin, ymax;   /* BINS_Y, YMIN, YMAX */
  int nz; double zmin, zmax;   /* BINS_Z, ZMIN, ZMAX */
  int proc_id;           /* PROC_ID: 11, 13, 15 (dileptons) */
  double lep_a;          /* LEP_A */
} upc_config;

/* Scalars derived at init (UpcCrossSection::init, src/UpcCrossSection.cpp:116-137). */
typedef struct upc_info {
  double g1, g2;         /* Lorentz factors */
  double factor;         /* Z^2 alpha / (pi^2 hc^2) */
  double mNucl;          /* nucleon mass */
  double rho0;           /* Woods-Saxon normalisation (calcWSRho) */
  double csNN;           /* sigma_NN [fm^2] used by G_AA (prepareGAA) */
  double p20;            /* breakup probability at the 20 fm clamp (1 for mode 1) */
  long long qags_calls;  /* form-factor flux integrals run by the last grid */
  long long qags_bad;    /* of which ended with ier != 0 (the reference aborts) */
  long long qags_evals;  /* integrand evaluations of the last grid */
  long long mfma_ksteps; /* 16-point x 16-column x 32-deep MFMA tile steps of the last MFMA grid */
  double mfma_spacing;   /* shared b-grid spacing of the MFMA forms [fm] */
  long long surrogate_points;    /* bins of the last surrogate grid served by the network */
  long long surrogate_fallback;  /* bins of the last surrogate grid recomputed by quadrature */
  double bskip;          /* [fm]: below it |G_AA spline| < p20 2^-70, so the FP64 contraction
                            takes P G as 0 there (no reference counterpart) */
  long long mfma_fp64_items; /* work items of the last MFMA grid that the cancellation guard
                                recomputed with the FP64 contraction (upc_ctx_set_mfma_cancel_limit) */
} upc_info;

/* Phases timed with HIP events when timing is enabled (upc_ctx_set_timing). */
#define UPC_PHASE_TABLES 0  /* init tables: rho0, G_AA, form factor, breakup */
#define UPC_PHASE_FLUX 1    /* flux rows: geometry, point-charge flux, task lists */
#define UPC_PHASE_QAGS 2    /* form-factor flux integrals (QAGS kernel) */
#define UPC_PHASE_LUMI 3    /* luminosity contraction (FP64 / MFMA / surrogate) */
#define UPC_PHASE_XSEC 4    /* elementary and nuclear cross-section tables */
#define UPC_PHASE_PDF 5     /* sampler CDF builds */
#define UPC_PHASE_PROJECT 6 /* MFMA forms: kernel matrix on the shared grid + flux-row projection */
#define UPC_NPHASES 8

typedef struct upc_timings {
  double ms[UPC_NPHASES];       /* summed device time per phase since the last reset */
  long long launches[UPC_NPHASES];
} upc_timings;

typedef struct upc_ctx upc_ctx;

const char *upc_last_error(void);
int upc_device_count(int *h_count);

/* UpcCrossSection::init minus the luminosity grid: calcWSRho, prepareGAA,
 * prepareFormFac, prepareBreakupProb (src/UpcCrossSection.cpp:116-131). All
 * tables are built on the GPU. */
int upc_ctx_create(const upc_config *cfg, void *stream, upc_ctx **out);
int upc_ctx_destroy(upc_ctx *ctx);
int upc_ctx_info(upc_ctx *ctx, upc_info *h_info);
/* Rebuild the tables of an existing context in place (no allocation). */
int upc_ctx_rebuild_tables(upc_ctx *ctx, void *stream);
/* Event timing of the phases (no reference counterpart: replaces its
 * progress logging, src/UpcCrossSection.cpp:554-557).  upc_ctx_timings
 * synchronises on the recorded events and resets the sums when reset != 0. */
int upc_ctx_set_timing(upc_ctx *ctx, int enable);
/* Node spacing h [fm] of the shared b grid of the MFMA forms (default 0.2;
 * accepted range [0.05, 2]).  No reference counterpart: the reference has
 * only the per-bin node quadrature. */
int upc_ctx_set_mfma_spacing(upc_ctx *ctx, double h_fm);
/* Cancellation guard of the MFMA forms (no reference counterpart).  A bin's
 * MFMA value is F + P + E: F = s0 sum(w1) sum(w2) (FP64, the phi sum with
 * every node beyond 20 fm), P the MFMA product of the flux rows with the
 * kernel's difference from s0, E the exact edge part.  P's error is a small
 * fraction (of order 1e-7) of |F|, so where P cancels most of F (flux
 * concentrated below 2R: high M at low sqrt(s)) the bin's relative error
 * grows with |F| / |L| (measured: about 2.5e-6 |F| / |L|).  Bins with
 * |F| > limit |L| (default 2, which keeps every bin of the tested systems
 * within 1e-5; for polarised grids either output) are recomputed by the
 * FP64 contraction from the same flux rows; limit <= 0 turns the guard
 * off.  upc_info's mfma_fp64_items counts them. */
int upc_ctx_set_mfma_cancel_limit(upc_ctx *ctx, double limit);
/* Intervals (1..16, default 16) a form-factor flux integral may hold in the
 * fast QAGS kernel before it is handed to the kernel with the reference's
 * 1000-interval workspace.  The result does not depend on it (both kernels
 * run the same arithmetic); smaller values exercise the second kernel.  No
 * reference counterpart. */
int upc_ctx_set_qags_cap(upc_ctx *ctx, int cap);
int upc_ctx_timings(upc_ctx *ctx, upc_timings *h_out, int reset);

/* ---- per-point functions of the path (batched; device arrays of n) ---- */
/* UpcCrossSection::fluxPoint(b, k)            src/UpcCrossSection.cpp:166-178 */
int upc_flux_point(upc_ctx *ctx, const double *d_b, const double *d_k, int n, double *d_out,
                   void *stream);
/* UpcCrossSection::fluxForm(b, k)             src/UpcCrossSection.cpp:194-218
 * d_ier (may be NULL) receives the QAGS status per point: dqagse's ier
 * (0 = converged, 1..5), 99 = still unfinished at the 1000-interval limit
 * (not reached by QUADPACK's own stopping rules), or 100 = a node of a QK21
 * pass lies past the last form-factor knot (t > 1.999998 GeV^2), where the
 * reference's gsl_spline_eval raises GSL_EDOM and its error handler ends the
 * program: the integral stops after that pass and its flux is NaN. */
int upc_flux_form(upc_ctx *ctx, const double *d_b, const double *d_k, int n, double *d_out,
                  int *d_ier, void *stream);
/* upc_flux_form without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as upc_flux_form; its work buffers are one file-scope
 * device array per device (3 MB, held until the process ends) instead of
 * the context's.  At most 262,144 points per call (UPC_EINVAL beyond).  Calls
 * that overlap on one device race on that array; the context's QAGS counters
 * (upc_ctx_info) are not updated.  The current device must be the
 * context's. */
int upc_flux_form_noalloc(upc_ctx *ctx, const double *d_b, const double *d_k, int n, double *d_out,
                          int *d_ier, void *stream);
/* J1(x), x >= 0, exactly as the form-factor flux integrand evaluates it on
 * the device (role of gsl_sf_bessel_J1 in fluxForm, src/UpcCrossSection.cpp:189).
 * No context needed. */
int upc_bessel_j1(const double *d_x, int n, double *d_out, void *stream);
/* UpcCrossSection::calcFormFac(Q2) (tabulated = 0) or its spline (1)
 *                                              src/UpcCrossSection.cpp:436-461 */
int upc_form_factor(upc_ctx *ctx, const double *d_q2, int n, int tabulated, double *d_out,
                    void *stream);
/* UpcCrossSection::calcBreakupProb(b, mode) (tabulated = 0) or the
 * prepareBreakupProb spline (1, b <= 20)       src/UpcCrossSection.cpp:416-434, 752-1019 */
int upc_breakup_prob(upc_ctx *ctx, const double *d_b, int n, int mode, int tabulated,
                     double *d_out, void *stream);
/* G_AA spline (which = 0) or T_A spline (which = 1) of prepareGAA
 *                                              src/UpcCrossSection.cpp:364-414 */
int upc_nuclear_overlap(upc_ctx *ctx, const double *d_b, int n, int which, double *d_out,
                        void *stream);
/* calcTwoPhotonLumi(M, Y) / calcTwoPhotonLumiPol(ns, np, M, Y) at n
 * arbitrary points                              src/UpcCrossSection.cpp:221-335
 * Unpolarised -> d_L; polarised -> d_Ls, d_Lp. */
int upc_lumi_points(upc_ctx *ctx, const double *d_M, const double *d_Y, int n, double *d_L,
                    double *d_Ls, double *d_Lp, void *stream);

/* UpcCrossSection::calcPhotonFlux(M, Y): photon flux at k = M/2 e^Y weighted
 * by P(min(b, 20)) G_AA(b) on the flux row's nodes (vector-meson path)
 *                                              src/UpcCrossSection.cpp:699-720 */
int upc_photon_flux(upc_ctx *ctx, const double *d_M, const double *d_Y, int n, double *d_out, void *stream);
/* UpcCrossSection::calcNucCrossSectionY(cs, ratio, totCS) for a particle of
 * mass mpart, with the elementary cross sections given at the ny lower bin
 * edges: d_cs_pos[iy] = sigma(y), d_cs_neg[iy] = sigma(-y) (the reference
 * takes them from UpcElemProcess::calcCrossSectionY).  d_cs[iy], d_ratio[iy];
 * *h_tot += sum, then the sum is multiplied by dy as in the reference.
 *                                              src/UpcCrossSection.cpp:722-748 */
int upc_nuc_cs_y(upc_ctx *ctx, double mpart, const double *d_cs_pos, const double *d_cs_neg, double *d_cs,
                 double *d_ratio, double *h_tot, void *stream);

/* ---- grids ---- */
/* prepareTwoPhotonLumi's grid: [nm][ny], lower bin edges, x dm x dy
 *                                              src/UpcCrossSection.cpp:494-576 */
int upc_lumi_grid(upc_ctx *ctx, int form, double *d_L, double *d_Ls, double *d_Lp, void *stream);
/* fillCrossSectionZM(cs, zmin, zmax, nz, mmin, mmax, nm, flag): d_zm [nm][nz]
 * at lower bin edges, x hc^2 1e7 / dm      src/UpcCrossSection.cpp:337-362 */
int upc_cs_zm(upc_ctx *ctx, double zmin, double zmax, int nz, double mmin, double mmax, int nm, int flag,
              double *d_zm, void *stream);
/* Elementary sigma(M) (flag 0), sigma_S(M) (1), sigma_PS(M) (2) at nm lower edges
 *                                              src/UpcTwoPhotonDilep.cpp:46-121 */
int upc_cs_m(upc_ctx *ctx, int flag, double *d_cs, void *stream);
/* UpcTwoPhotonDilep::calcCrossSectionM/MPolS/MPolPS (flag 0/1/2; d_z ignored)
 * or calcCrossSectionZM/ZMPolS/ZMPolPS (flag 3/4/5) at n points  src/UpcTwoPhotonDilep.cpp:46-134 */
int upc_elem_cs(upc_ctx *ctx, int flag, const double *d_m, const double *d_z, int n, double *d_out,
                void *stream);
/* calcNucCrossSectionYM(cs, ratio, totCS): cs/ratio [ny][nm]; *h_tot_cs is
 * incremented by the total in mb like the reference   src/UpcCrossSection.cpp:594-698 */
int upc_nuc_cs_ym(upc_ctx *ctx, const double *d_L, const double *d_Ls, const double *d_Lp,
                  double *d_cs, double *d_ratio, double *h_tot_cs, void *stream);

/* ---- sampler tables (include/UpcSampler.h:37-142) ---- */
/* gsl_histogram_pdf_init on `nrows` independent rows of `n` bins:
 * d_sum[r*(n+1) + i].  Returns UPC_EINVAL if any bin is negative (GSL_EDOM). */
int upc_pdf_init(const double *d_bins, int nrows, int n, double *d_sum, void *stream);
/* Draw n events (Y, M, z) from the generator's tables like
 * UpcGenerator::generateEvent (src/UpcGenerator.cpp:735-757): the 2-D (Y,M)
 * pdf d_sum_ym [(ny*nm)+1], the per-M z pdfs d_sum_z [nm][nz+1] (polarised:
 * d_sum_zs and d_sum_zp with d_ratio [ny][nm]).  Counter-based Philox4x32-10,
 * stream (seed, offset). */
int upc_sample_events(upc_ctx *ctx, const double *d_sum_ym, const double *d_sum_z,
                      const double *d_sum_zs, const double *d_sum_zp, const double *d_ratio,
                      unsigned long long seed, unsigned long long offset, long long n,
                      double *d_y, double *d_m, double *d_z, void *stream);

/* ---- event generation for the dilepton processes ---- */
/* Kinematic options of UpcGenerator (parameters.in keys NON_ZERO_GAM_PT,
 * DO_PT_CUT, PT_MIN, DO_ETA_CUT, ETA_MIN, ETA_MAX). */
typedef struct upc_event_opts {
  int nonzero_gam_pt;
  int do_pt_cut;
  double pt_min;
  int do_eta_cut;
  double eta_min, eta_max;
} upc_event_opts;

/* Photon transverse-momentum sampling tables of getPhotonPt, built from the
 * context's form-factor table for Lorentz factor gtot
 *                                              src/UpcCrossSection.cpp:1021-1051 */
int upc_photon_pt_tables(upc_ctx *ctx, double gtot, void *stream);
/* n getPhotonPt(E) draws at photon energies d_E (stream seed). */
int upc_photon_pt(upc_ctx *ctx, const double *d_E, long long n, unsigned long long seed, double *d_out,
                  void *stream);
/* n pair events as UpcGenerator::generateEvent (src/UpcGenerator.cpp:715-790):
 * (Y, M) and z as upc_sample_events, pair momentum (getPairMomentum), lepton
 * 4-momenta (pairProduction) and checkKinCuts.  d_p [n][2][4] = (px, py, pz,
 * E) per lepton, d_pdg [n][2], d_pass[n] = 1 accepted, 2 rejected by cuts,
 * 0 not drawn (random number beyond the last CDF entry). */
int upc_generate_pairs(upc_ctx *ctx, const double *d_sum_ym, const double *d_sum_z, const double *d_sum_zs,
                       const double *d_sum_zp, const double *d_ratio, const upc_event_opts *opts,
                       unsigned long long seed, unsigned long long offset, long long n, double *d_p, int *d_pdg,
                       int *d_pass, void *stream);

/* ---- batches: many configurations in one launch sequence ----
 * No reference counterpart as one call: a batch of n jobs computes what n
 * reference processes compute in UpcCrossSection::init + prepareTwoPhotonLumi
 * (src/UpcCrossSection.cpp:116-137, 463-592), each job with its own system,
 * breakup mode, flux type, polarisation and grid.  The flux integrals of all
 * jobs share one persistent QAGS launch and the contraction is one launch;
 * upc_batch_run(use_graph = 1) captures the sequence into a hipGraph on its
 * first call and replays it afterwards. */
typedef struct upc_batch upc_batch;

typedef struct upc_tune {
  int qags_blocks_per_cu;     /* chosen QAGS workgroups (256 threads) per CU */
  int lumi_waves;             /* chosen contraction waves per workgroup */
  int qags_candidates[3];
  double qags_ms[3];          /* QAGS kernel time of each candidate */
  int lumi_candidates[3];
  double lumi_ms[3];          /* FP64 contraction time of each candidate (FP64 form) */
  int form;                   /* the batch's contraction form when tuned */
  int mfma_tile_group;        /* chosen MFMA column tiles per A-fragment load (MFMA forms) */
  int mfma_candidates[3];
  double mfma_ms[3];          /* MFMA contraction time (every job) of each candidate (MFMA forms) */
} upc_tune;

typedef struct upc_batch_stats {
  int njobs;
  long long points, rows;
  long long qags_calls, qags_evals, qags_bad;   /* of the last run */
  int qags_blocks_per_cu, lumi_waves;           /* launch shapes in use */
  long long tasks;        /* form-factor flux integrals per run (counted at creation) */
  long long work_bytes;   /* device memory of the batch's work buffers (the jobs' tables not included) */
  int form;               /* contraction form (upc_batch_set_form) */
  int mfma_tile_group;    /* MFMA tile shape in use */
  long long mfma_fp64_items; /* MFMA forms: work items of the last run recomputed in FP64 by the
                                cancellation guard (upc_ctx_set_mfma_cancel_limit, default limit) */
} upc_batch_stats;

/* Builds one context (tables) per job and the batch work buffers. */
int upc_batch_create(const upc_config *cfgs, int njobs, void *stream, upc_batch **out);
int upc_batch_destroy(upc_batch *batch);
/* Output grids per job: d_L[j] (unpolarised jobs) or d_Ls[j], d_Lp[j]
 * (polarised), [nm][ny] x dm x dy like upc_lumi_grid.  Arrays of n device
 * pointers; unused entries may be NULL.  Captured graphs see new pointers. */
int upc_batch_set_outputs(upc_batch *batch, double *const *d_L, double *const *d_Ls, double *const *d_Lp);
/* FP64 grids of every job.  use_graph: capture once, then replay. */
int upc_batch_run(upc_batch *batch, int use_graph, void *stream);
/* Times every candidate launch shape on this batch and keeps the fastest
 * (also computes the grids).  Synchronises `stream`. */
int upc_batch_autotune(upc_batch *batch, void *stream, upc_tune *h_out);
/* Sets the launch shapes directly (values from upc_tune's candidate lists). */
int upc_batch_set_tune(upc_batch *batch, int qags_blocks_per_cu, int lumi_waves);
/* Contraction form of every job: UPC_FORM_FP64 (default, the reference's
 * quadrature), UPC_FORM_MFMA or UPC_FORM_MFMA_FP32 (FP64 flux rows, the
 * shared-grid contraction of upc_lumi_grid's MFMA forms: each job's grid is
 * the one its own context computes in that form).  The MFMA forms allocate
 * one shared-grid state per polarisation here, not at run time. */
int upc_batch_set_form(upc_batch *batch, int form);
/* MFMA tile shape: column tiles that share one A-fragment load (1, 2 or 4;
 * results are identical, only the time differs). */
int upc_batch_set_mfma_tiles(upc_batch *batch, int tile_group);
int upc_batch_info(upc_batch *batch, void *stream, upc_batch_stats *h_out);
/* Per job of the last run: form-factor flux integrals that ended with a QAGS
 * error or a form-factor spline domain error (h_failed[njobs]).  The
 * reference process of that configuration stops at the first of them (GSL's
 * default error handler, src/UpcCrossSection.cpp:209 and :188); the library
 * keeps QUADPACK's value, NaN for a domain error.  Synchronises `stream`. */
int upc_batch_job_failures(upc_batch *batch, void *stream, long long *h_failed);

/* ---- learned surrogate of log L(M, Y; system) ----
 * No reference counterpart (the reference always integrates).  Two MLPs
 * (inputs: log M, Y, log gamma, Z, A^(1/3), R, a, one-hot breakup mode,
 * log xi1, log xi2 with xi = k R / (gamma hc); 3 x
 * 128 SiLU; bf16 operands, FP32 accumulation on MFMA) predict
 * r = log L - log L0 with L0 = (M/2) n(k1) n(k2) the analytic point-charge
 * luminosity beyond b = R; targets are normalised (r - mu) / sigma.
 * Trained by the library's own forward/backward/Adam kernels. */
typedef struct upc_sur upc_sur;
int upc_sur_create(unsigned long long seed, upc_sur **out);
int upc_sur_destroy(upc_sur *sur);
int upc_sur_param_count(int *h_n);   /* floats per network */
int upc_sur_get_params(upc_sur *sur, int net, float *h_out);
int upc_sur_set_params(upc_sur *sur, int net, const float *h_in);
int upc_sur_set_norm(upc_sur *sur, double mu, double sigma);
int upc_sur_get_norm(upc_sur *sur, double *h_mu, double *h_sigma);
/* Product precision of both networks (forward, gradients, inference):
 * split = 1 (default) splits every operand into bf16 hi + lo and sums
 * hi*hi + hi*lo + lo*hi on MFMA with FP32 accumulation (about FP32 products),
 * the output layer uses the FP32 weights; split = 0 rounds each operand to
 * one bf16 (and the output weights too). */
int upc_sur_set_precision(upc_sur *sur, int split);
/* Predicted error of a bin = scale * |r_a - r_b| (the ensemble spread scaled
 * by a factor calibrated on held-out data; default 1). */
int upc_sur_set_err_scale(upc_sur *sur, double scale);
int upc_sur_get_err_scale(upc_sur *sur, double *h_scale);
/* Features [n][32] (FP32) and log L0 of n (M, Y) points of one system.
 * log L0 is NaN outside the surrogate's domain, max(xi1, xi2) <= 4 with
 * xi = k R / (gamma hc): there L is exponentially small and the QAGS values
 * are mostly quadrature error, so such bins are always integrated. */
int upc_sur_features(const upc_config *cfg, const double *d_M, const double *d_Y, long long n, float *d_X,
                     double *d_logL0, void *stream);
/* Normalised prediction of network `net` for B rows (B a multiple of 32, <= 65536). */
int upc_sur_forward(upc_sur *sur, int net, const float *d_X, int B, float *d_pred, void *stream);
/* Gradient of mean((pred - t)^2) w.r.t. all parameters of `net` (param_count floats). */
int upc_sur_grad(upc_sur *sur, int net, const float *d_X, const float *d_T, int B, float *d_grad, void *stream);
/* Adam training of both networks on n rows (mini-batches drawn with Philox),
 * from their current weights with fresh moments (so a second call fine-tunes);
 * learning rate lr decaying by a cosine to lr/100 over `steps`.  use_graph:
 * one captured step replayed `steps` times.  h_loss (optional): [2][steps]
 * mean squared error per step. */
int upc_sur_train(upc_sur *sur, const float *d_X, const float *d_T, long long n, int steps, int batch, double lr,
                  int use_graph, void *stream, float *h_loss);
/* Attach a trained surrogate to a context for upc_lumi_grid(UPC_FORM_SURROGATE):
 * bins whose predicted error exceeds max_err (in log L) are
 * recomputed by the FP64 quadrature (the spread times the surrogate's error
 * scale), as are bins outside its domain.  sur = NULL detaches. */
int upc_ctx_set_surrogate(upc_ctx *ctx, upc_sur *sur, double max_err);
/* Error check of the surrogate grid form (no reference counterpart).  With
 * anchor spacings every_m, every_y > 0 (default 16 and 4) the FP64 quadrature
 * is run at the anchor lattice (every every_m-th M bin x every every_y-th Y
 * bin, and the last ones), anchor bins take their quadrature values, and any
 * other bin takes the network's value only when the network's |d log L| at
 * all four anchors around it is within max_err / 2; the rest, and every bin
 * outside the network's domain, fall back to the quadrature.  0, 0 selects
 * the ensemble-spread gate instead (err_scale * |r_a - r_b| <= max_err). */
int upc_ctx_set_surrogate_anchors(upc_ctx *ctx, int every_m, int every_y);

#ifdef __cplusplus
}
#endif
#endif /* UPC_LUMI_H */
