/*
 * upc_oracle.h -- CPU FP64 restatement of upcgen's luminosity / nuclear
 * cross-section path.  TEST INFRASTRUCTURE ONLY: this library is the parity
 * checker for the HIP product and the timed CPU baseline.  Nothing in the
 * product (upcgen_amd/) may include, link or call it.
 *
 * Every function cites the reference file:line it restates.  Reference paths
 * are relative to the upcgen source tree (src/, include/).
 */
#ifndef UPC_ORACLE_H
#define UPC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Physics + grid configuration: the parameters.in keys of
 * src/UpcGenerator.cpp:183-305 that reach UpcCrossSection. */
typedef struct orc_params {
  int Z, A;              /* NUCLEUS_Z, NUCLEUS_A */
  double R, a;           /* WS_R, WS_A (fm) */
  double sqrts;          /* SQRTS (GeV); g1 = g2 = sqrts / (2 m_p) */
  int flux_point;        /* FLUX_POINT */
  int breakup_mode;      /* BREAKUP_MODE 1..4 */
  int use_polarized;     /* USE_POLARIZED_CS */
  int nm; double mmin, mmax;
  int ny; double ymin, ymax;
  int nz; double zmin, zmax;
  int proc_id;           /* 11, 13, 15 (dileptons) */
  double lep_a;          /* LEP_A */
  int num_threads;       /* -nthreads */
} orc_params;

typedef struct orc_ctx orc_ctx;

/* ---- special functions and numerics (unit-tested against KAT fixtures) ---- */
double orc_bessel_k0(double x);        /* accurate K0 (gsl_sf_bessel_K0 role) */
double orc_bessel_k1(double x);        /* accurate K1 (gsl_sf_bessel_K1 role) */
double orc_bessel_j1(double x);        /* accurate J1 (gsl_sf_bessel_J1 role) */
double orc_bessel_k1_as(double x);     /* A&S 9.8.7/9.8.8 K1 (TMath::BesselK1 role) */
double orc_bessel_i1_as(double x);     /* A&S 9.8.3/9.8.4 I1 (TMath::BesselI1 role) */
double orc_simpson(int n, const double *v, double h);

typedef double (*orc_fn)(double x, void *p);
/* ier of a form-factor flux integral stopped at its first QK21 pass with a
 * NaN value: a node past the last form-factor knot, where the reference's
 * gsl_spline_eval raises GSL_EDOM (its default handler ends the program). */
#define ORC_QAGS_EDOM 100
/* QUADPACK dqagse (QK21 + Wynn epsilon).  Returns ier (0 = success). */
int orc_qags(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
             int limit, double *result, double *abserr, int *neval, int *last);
/* Same, and also copies the final interval list (alist, blist, rlist, elist in
 * storage order) into caller arrays of length >= limit. */
int orc_qags_trace(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
                   int limit, double *result, double *abserr, int *neval, int *last,
                   double *alist, double *blist, double *rlist, double *elist);

/* Natural cubic spline with gsl_interp_cspline semantics. */
typedef struct orc_spline { int n; double *x, *y, *c; } orc_spline;
int orc_spline_init(orc_spline *s, const double *x, const double *y, int n);
double orc_spline_eval(const orc_spline *s, double x, int *err);
void orc_spline_free(orc_spline *s);

/* ---- context: UpcCrossSection::init (src/UpcCrossSection.cpp:116-137) ---- */
orc_ctx *orc_create(const orc_params *p);   /* rho0, G_AA, FF and breakup tables */
void orc_destroy(orc_ctx *c);
/* Wall seconds orc_create spent on the tables: out[0] rho0 + G_AA, out[1]
 * the form-factor spline, out[2] the breakup tables and spline. */
void orc_table_times(const orc_ctx *c, double *out);

void orc_set_qags_tol(orc_ctx *c, double epsabs, double epsrel);
double orc_rho0(const orc_ctx *c);
double orc_factor(const orc_ctx *c);
double orc_g1(const orc_ctx *c);
double orc_gaa(const orc_ctx *c, double b);             /* G_AA spline */
double orc_ta(const orc_ctx *c, double b);              /* T_A spline */
double orc_cs_nn(const orc_ctx *c);
double orc_formfac(const orc_ctx *c, double Q2);        /* analytic F(Q2) */
double orc_formfac_spline(const orc_ctx *c, double Q2); /* tabulated F */
double orc_breakup_prob(const orc_ctx *c, double b, int mode); /* Starlight sum */
double orc_breakup_spline(const orc_ctx *c, double b);  /* tabulated P(b) */
/* The XN sum p_xn(b) of calcBreakupProb (src/UpcCrossSection.cpp:972-992);
 * -1 where the reference returns 0 before summing. */
double orc_breakup_pxn(const orc_ctx *c, double b);
/* calcBreakupProb's first-call tables (src/UpcCrossSection.cpp:852-969):
 * ee[1..n], se[1..n] (n returned; ee[n] is the 99999999999 sentinel; arrays
 * of >= 10002), eee[1..161], sa[1..160] (>= 162), zcon and o0. */
int orc_breakup_tables(const orc_ctx *c, double *ee, double *se, double *eee, double *sa, double *zcon,
                       double *o0);
/* fluxFormIntegrand(q) at (b, k) (src/UpcCrossSection.cpp:181-191). */
double orc_flux_integrand(const orc_ctx *c, double b, double k, double q);
double orc_flux_point(const orc_ctx *c, double b, double k);
double orc_flux_form(const orc_ctx *c, double b, double k, int *neval, int *ier);
double orc_lumi(const orc_ctx *c, double M, double Y);
void orc_lumi_pol(const orc_ctx *c, double M, double Y, double *ns, double *np);
/* prepareTwoPhotonLumi grid, row-major [im][iy], already x dm x dy.
 * out (unpolarised) or outS/outP (polarised).  Returns the number of QAGS
 * calls that ended with ier != 0 (the reference would abort on any). */
long orc_lumi_grid(const orc_ctx *c, double *out, double *outS, double *outP);
/* The same grid; counts[(im*ny+iy)*4 + 0..3] (may be NULL) receive the QAGS
 * integrand evaluations of the bin's k1 and k2 flux rows (120 fluxForm calls
 * each, src/UpcCrossSection.cpp:235-269) and how many of those integrals
 * ended with ier != 0 (k1, k2). */
long orc_lumi_grid_counts(const orc_ctx *c, double *out, double *outS, double *outP, long *counts);
/* calcPhotonFlux(M, Y) and calcNucCrossSectionY (src/UpcCrossSection.cpp:699-748). */
double orc_photon_flux(const orc_ctx *c, double M, double Y);
void orc_nuc_cs_y(const orc_ctx *c, double mpart, const double *cs_pos, const double *cs_neg, double *cs,
                  double *ratio, double *tot_cs);

/* Elementary dilepton cross sections, src/UpcTwoPhotonDilep.cpp:46-134. */
double orc_dilep_cs_m(int pdg, double lep_a, double m);
double orc_dilep_cs_zm(int pdg, double lep_a, double z, double m);
double orc_dilep_cs_m_pol_s(int pdg, double m);
double orc_dilep_cs_zm_pol_s(int pdg, double z, double m);
double orc_dilep_cs_m_pol_ps(int pdg, double m);
double orc_dilep_cs_zm_pol_ps(int pdg, double z, double m);

/* fillCrossSectionZM, out[im][iz]. */
void orc_fill_cs_zm(const orc_ctx *c, double *out, int flag);
/* calcNucCrossSectionYM: cs[iy][im], ratio[iy][im]; lumi grids [im][iy].
 * *tot_cs accumulates like the reference (caller zeroes it), then x1e-6. */
void orc_nuc_cs_ym(const orc_ctx *c, const double *lumi, const double *lumiS,
                   const double *lumiP, double *cs, double *ratio, double *tot_cs);

/* getPhotonPt histogram contents (5000 bins over (0, 6 hc/R]) at energy E. */
void orc_photon_pt_hist(const orc_ctx *c, double E, double gtot, double *out);

/* gsl_histogram_pdf_init semantics: sum[0..n] from n non-negative bins.
 * Returns 0, or -1 if a bin is negative (GSL_EDOM). */
int orc_pdf_init(const double *bins, int n, double *sum);
/* gsl_histogram_pdf_sample: inverse CDF with linear placement in the bin. */
double orc_pdf_sample(const double *sum, const double *range, int n, double r, int *bin);

/* mt19937 with gsl_rng_set / gsl_rng_uniform semantics. */
typedef struct orc_mt { unsigned long mt[624]; int mti; } orc_mt;
void orc_mt_seed(orc_mt *s, unsigned long seed);
double orc_mt_uniform(orc_mt *s);

#ifdef __cplusplus
}
#endif
#endif
