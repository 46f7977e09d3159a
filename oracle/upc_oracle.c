/*
 * upc_oracle.c -- CPU FP64 restatement of upcgen's two-photon luminosity and
 * nuclear cross-section path (TEST INFRASTRUCTURE: the parity checker and the
 * timed CPU baseline; never linked into the product).
 *
 * Reference anchors (upcgen source tree):
 *   src/UpcCrossSection.cpp:116-137   init
 *   src/UpcCrossSection.cpp:139-163   simpson, calcWSRho
 *   src/UpcCrossSection.cpp:165-218   fluxPoint, fluxFormIntegrand, fluxForm
 *   src/UpcCrossSection.cpp:220-335   calcTwoPhotonLumi, calcTwoPhotonLumiPol
 *   src/UpcCrossSection.cpp:337-362   fillCrossSectionZM
 *   src/UpcCrossSection.cpp:364-461   prepareGAA, prepareBreakupProb, calcFormFac, prepareFormFac
 *   src/UpcCrossSection.cpp:463-592   prepareTwoPhotonLumi (grid loop, OpenMP row split)
 *   src/UpcCrossSection.cpp:594-698   calcNucCrossSectionYM
 *   src/UpcCrossSection.cpp:752-1019  calcBreakupProb (Starlight)
 *   src/UpcTwoPhotonDilep.cpp:29-134  dilepton elementary cross sections
 *   include/UpcSampler.h:37-142       histogram-pdf samplers
 *   include/UpcPhysConstants.h:24-33  constants
 *
 * Third-party numerics the reference calls, restated from published methods:
 *   gsl_sf_bessel_K0/K1 -> DLMF 10.31.1 power series (x <= 1) and the
 *                          trapezoidal rule on DLMF 10.32.9 (x > 1)
 *   gsl_sf_bessel_J1    -> C library j1()
 *   TMath::BesselK1/I1  -> Abramowitz & Stegun 9.8.3, 9.8.4, 9.8.7, 9.8.8
 *   gsl_integration_qags-> QUADPACK dqagse/dqk21/dqpsrt/dqelg (Piessens et al. 1983)
 *   gsl_interp_cspline  -> natural cubic spline, LDL^T tridiagonal solve
 *   gsl_histogram_pdf   -> running-mean normalised prefix sums, inverse CDF
 *   gsl_rng_mt19937     -> Matsumoto & Nishimura 1998, 2002 seeding
 */
#define _GNU_SOURCE
#include "upc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* include/UpcPhysConstants.h:26-32 */
static const double kAlpha = 1.0 / 137.035999074;
static const double kHc = 0.1973269718;
static const double kMProt = 0.9382720813;
static const double kMNeut = 0.939565346;
static const double kMEl = 0.000510998946;
static const double kMMu = 0.1056583745;
static const double kMTau = 1.77686;

/* include/UpcCrossSection.h:90-111: 10-point Gauss-Legendre table as printed
 * there; only the 5 non-negative abscissas are used by every caller. */
static const double kGLw[10] = {0.0666713443086881, 0.1494513491505806, 0.2190863625159820,
                                0.2692667193099963, 0.2955242247147529, 0.2955242247147529,
                                0.2692667193099963, 0.2190863625159820, 0.1494513491505806,
                                0.0666713443086881};
static const double kGLx[10] = {-0.9739065285171717, -0.8650633666889845, -0.6794095682990244,
                                -0.4333953941292472, -0.1488743389816312, 0.1488743389816312,
                                0.4333953941292472, 0.6794095682990244, 0.8650633666889845,
                                0.9739065285171717};

/* src/UpcCrossSection.cpp:46-49 */
static const double kQ2min = 1e-9;
static const double kQ2max = 2.;
static const int kNQ2 = 1000000;
#define DQ2 ((kQ2max - kQ2min) / kNQ2)

/* include/UpcCrossSection.h:116-117, 139 */
enum { NB1 = 120, NB2 = 120, NB = 200 };

/* ------------------------------------------------------------------------ */
/* Special functions                                                         */
/* ------------------------------------------------------------------------ */
static const double kEuler = 0.57721566490153286060651209008240243;

/* DLMF 10.31.1 (n = 0, 1) for 0 < x <= 1. */
static void k01_series(double x, double *k0, double *k1)
{
  const double y = 0.25 * x * x;
  const double lx = log(0.5 * x);
  double t = 1.0, u = 1.0;       /* t_k = y^k/(k!)^2, u_k = y^k/(k!(k+1)!) */
  double i0 = 1.0, s0 = 0.0;     /* I0 and sum H_k t_k */
  double i1s = 1.0;              /* I1 / (x/2) */
  double s1 = 1.0 - 2.0 * kEuler; /* sum (psi(k+1)+psi(k+2)) u_k, k=0 term */
  double hk = 0.0;
  for (int k = 1; k < 30; ++k) {
    t *= y / ((double)k * k);
    u *= y / ((double)k * (k + 1));
    hk += 1.0 / k;
    i0 += t;
    s0 += hk * t;
    i1s += u;
    s1 += (2.0 * hk + 1.0 / (k + 1) - 2.0 * kEuler) * u;
    if (t < 1e-20 * i0 && u < 1e-20) break;
  }
  *k0 = -(lx + kEuler) * i0 + s0;
  *k1 = 1.0 / x + lx * (0.5 * x * i1s) - 0.25 * x * s1;
}

/* Trapezoidal rule on K_nu(x) e^x = int_0^inf exp(-2x sinh^2(t/2)) cosh(nu t) dt
 * (DLMF 10.32.9), whose integrand is entire: the error falls like
 * exp(-pi^2/h) for small x and exp(-2 pi^2/(h^2 x)) for large x. */
static void k01_scaled_trap(double x, double *k0s, double *k1s)
{
  double h = M_PI * M_PI / (50.0 + x);
  const double h2 = 0.6 / sqrt(x);
  if (h2 < h) h = h2;
  double s0 = 0.5, s1 = 0.5;
  for (int k = 1; k < 100000; ++k) {
    const double t = k * h;
    const double sh = sinh(0.5 * t);
    const double e = exp(-2.0 * x * sh * sh);
    s0 += e;
    s1 += e * cosh(t);
    if (e < 1e-19 * s0) break;
  }
  *k0s = h * s0;
  *k1s = h * s1;
}

double orc_bessel_k0(double x)
{
  if (x <= 0) return NAN;
  if (x <= 1.0) {
    double k0, k1;
    k01_series(x, &k0, &k1);
    return k0;
  }
  double k0s, k1s;
  k01_scaled_trap(x, &k0s, &k1s);
  return exp(-x) * k0s;
}

double orc_bessel_k1(double x)
{
  if (x <= 0) return NAN;
  if (x <= 1.0) {
    double k0, k1;
    k01_series(x, &k0, &k1);
    return k1;
  }
  double k0s, k1s;
  k01_scaled_trap(x, &k0s, &k1s);
  return exp(-x) * k1s;
}

double orc_bessel_j1(double x) { return j1(x); }

/* Abramowitz & Stegun 9.8.3 / 9.8.4 (the polynomial TMath::BesselI1 uses). */
double orc_bessel_i1_as(double x)
{
  const double ax = fabs(x);
  if (ax < 3.75) {
    const double t = x / 3.75;
    const double y = t * t;
    return x * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 +
               y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  }
  const double y = 3.75 / ax;
  double r = (exp(ax) / sqrt(ax)) *
             (0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 +
              y * (-0.1031555e-1 + y * (0.2282967e-1 + y * (-0.2895312e-1 +
              y * (0.1787654e-1 + y * -0.420059e-2))))))));
  return x < 0 ? -r : r;
}

/* Abramowitz & Stegun 9.8.7 / 9.8.8 (the polynomial TMath::BesselK1 uses). */
double orc_bessel_k1_as(double x)
{
  if (x <= 0) return 0.;
  if (x <= 2) {
    const double y = x * x / 4;
    return (log(x / 2) * orc_bessel_i1_as(x)) +
           (1. / x) * (1. + y * (0.15443144 + y * (-0.67278579 + y * (-0.18156897 +
            y * (-0.1919402e-1 + y * (-0.110404e-2 + y * (-0.4686e-4)))))));
  }
  const double y = 2 / x;
  return (exp(-x) / sqrt(x)) *
         (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 + y * (0.1504268e-1 +
          y * (-0.780353e-2 + y * (0.325614e-2 + y * (-0.68245e-3)))))));
}

/* src/UpcCrossSection.cpp:139-150: with even n the closing weights are 2, 1
 * (not a textbook Simpson rule); kept as written. */
double orc_simpson(int n, const double *v, double h)
{
  double sum = v[0] + v[n - 1];
  for (int i = 1; i < n - 1; i += 2) sum += 4. * v[i];
  for (int i = 2; i < n - 1; i += 2) sum += 2. * v[i];
  return sum * h / 3.;
}

/* ------------------------------------------------------------------------ */
/* Natural cubic spline (gsl_interp_cspline)                                 */
/* ------------------------------------------------------------------------ */
int orc_spline_init(orc_spline *s, const double *x, const double *y, int n)
{
  s->n = n;
  s->x = (double *)malloc(sizeof(double) * n);
  s->y = (double *)malloc(sizeof(double) * n);
  s->c = (double *)calloc(n, sizeof(double));
  memcpy(s->x, x, sizeof(double) * n);
  memcpy(s->y, y, sizeof(double) * n);
  const int m = n - 2; /* interior unknowns c[1..n-2]; c[0] = c[n-1] = 0 */
  if (m <= 0) return 0;
  double *diag = (double *)malloc(sizeof(double) * m);
  double *off = (double *)malloc(sizeof(double) * m);
  double *g = (double *)malloc(sizeof(double) * m);
  for (int i = 0; i < m; ++i) {
    const double hi = x[i + 1] - x[i];
    const double hip1 = x[i + 2] - x[i + 1];
    const double ydi = y[i + 1] - y[i];
    const double ydip1 = y[i + 2] - y[i + 1];
    const double gi = hi != 0.0 ? 1.0 / hi : 0.0;
    const double gip1 = hip1 != 0.0 ? 1.0 / hip1 : 0.0;
    off[i] = hip1;
    diag[i] = 2.0 * (hip1 + hi);
    g[i] = 3.0 * (ydip1 * gip1 - ydi * gi);
  }
  /* LDL^T of the symmetric tridiagonal matrix, forward and back substitution */
  double *gam = (double *)malloc(sizeof(double) * m);
  double *alp = (double *)malloc(sizeof(double) * m);
  double *z = (double *)malloc(sizeof(double) * m);
  alp[0] = diag[0];
  gam[0] = off[0] / alp[0];
  for (int i = 1; i < m - 1; ++i) {
    alp[i] = diag[i] - off[i - 1] * gam[i - 1];
    gam[i] = off[i] / alp[i];
  }
  if (m > 1) alp[m - 1] = diag[m - 1] - off[m - 2] * gam[m - 2];
  z[0] = g[0];
  for (int i = 1; i < m; ++i) z[i] = g[i] - gam[i - 1] * z[i - 1];
  for (int i = 0; i < m; ++i) z[i] = z[i] / alp[i];
  s->c[m] = z[m - 1];
  for (int i = m - 2; i >= 0; --i) s->c[i + 1] = z[i] - gam[i] * s->c[i + 2];
  free(diag); free(off); free(g); free(gam); free(alp); free(z);
  return 0;
}

void orc_spline_free(orc_spline *s)
{
  free(s->x); free(s->y); free(s->c);
  s->x = s->y = s->c = NULL;
}

/* gsl_interp_bsearch over [0, n-1]: x[i] <= v < x[i+1], i in [0, n-2]. */
static int bsearch_idx(const double *x, double v, int lo, int hi)
{
  while (hi > lo + 1) {
    const int i = (hi + lo) / 2;
    if (x[i] > v) hi = i; else lo = i;
  }
  return lo;
}

double orc_spline_eval(const orc_spline *s, double v, int *err)
{
  if (v < s->x[0] || v > s->x[s->n - 1]) { /* GSL_EDOM in the reference */
    if (err) *err = 1;
    return NAN;
  }
  const int i = bsearch_idx(s->x, v, 0, s->n - 1);
  const double xlo = s->x[i], xhi = s->x[i + 1];
  const double h = xhi - xlo;
  const double dy = s->y[i + 1] - s->y[i];
  const double dx = v - xlo;
  const double ci = s->c[i], cip1 = s->c[i + 1];
  const double bi = dy / h - h * (cip1 + 2.0 * ci) / 3.0;
  const double di = (cip1 - ci) / (3.0 * h);
  return s->y[i] + dx * (bi + dx * (ci + dx * di));
}

/* ------------------------------------------------------------------------ */
/* QUADPACK dqagse (QK21 + Wynn epsilon), 1-based internal arrays            */
/* ------------------------------------------------------------------------ */
static const double kXgk[12] = {0,
  0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
  0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
  0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
  0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
  0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
  0.000000000000000000000000000000000};
static const double kWgk[12] = {0,
  0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
  0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
  0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
  0.123491976262065851077208980748500, 0.134709217311473325928054001771707,
  0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
  0.149445554002916905664936468389821};
static const double kWg[6] = {0,
  0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
  0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
  0.295524224714752870173892994651338};

static void qk21(orc_fn f, void *p, double a, double b, double *result, double *abserr,
                 double *resabs, double *resasc)
{
  const double epmach = DBL_EPSILON, uflow = DBL_MIN;
  double fv1[11], fv2[11];
  const double centr = 0.5 * (a + b);
  const double hlgth = 0.5 * (b - a);
  const double dhlgth = fabs(hlgth);
  double resg = 0.0;
  const double fc = f(centr, p);
  double resk = kWgk[11] * fc;
  double rabs = fabs(resk);
  for (int j = 1; j <= 5; ++j) {
    const int jtw = 2 * j;
    const double absc = hlgth * kXgk[jtw];
    const double f1 = f(centr - absc, p);
    const double f2 = f(centr + absc, p);
    fv1[jtw] = f1; fv2[jtw] = f2;
    const double fsum = f1 + f2;
    resg = resg + kWg[j] * fsum;
    resk = resk + kWgk[jtw] * fsum;
    rabs = rabs + kWgk[jtw] * (fabs(f1) + fabs(f2));
  }
  for (int j = 1; j <= 5; ++j) {
    const int jtwm1 = 2 * j - 1;
    const double absc = hlgth * kXgk[jtwm1];
    const double f1 = f(centr - absc, p);
    const double f2 = f(centr + absc, p);
    fv1[jtwm1] = f1; fv2[jtwm1] = f2;
    const double fsum = f1 + f2;
    resk = resk + kWgk[jtwm1] * fsum;
    rabs = rabs + kWgk[jtwm1] * (fabs(f1) + fabs(f2));
  }
  const double reskh = resk * 0.5;
  double rasc = kWgk[11] * fabs(fc - reskh);
  for (int j = 1; j <= 10; ++j) rasc = rasc + kWgk[j] * (fabs(fv1[j] - reskh) + fabs(fv2[j] - reskh));
  *result = resk * hlgth;
  rabs = rabs * dhlgth;
  rasc = rasc * dhlgth;
  double err = fabs((resk - resg) * hlgth);
  if (rasc != 0.0 && err != 0.0) {
    const double sc = pow(200.0 * err / rasc, 1.5);
    err = rasc * (sc < 1.0 ? sc : 1.0);
  }
  if (rabs > uflow / (50.0 * epmach)) {
    const double m = (epmach * 50.0) * rabs;
    if (m > err) err = m;
  }
  *abserr = err;
  *resabs = rabs;
  *resasc = rasc;
}

static void qpsrt(int limit, int last, int *maxerr, double *ermax, const double *elist,
                  int *iord, int *nrmax)
{
  int i, k, isucc, jupbn, jbnd, ibeg;
  double errmax, errmin;
  if (last <= 2) {
    iord[1] = 1;
    iord[2] = 2;
    goto L90;
  }
  errmax = elist[*maxerr];
  if (*nrmax != 1) {
    const int ido = *nrmax - 1;
    for (i = 1; i <= ido; ++i) {
      isucc = iord[*nrmax - 1];
      if (errmax <= elist[isucc]) break;
      iord[*nrmax] = isucc;
      (*nrmax)--;
    }
  }
  jupbn = last;
  if (last > (limit / 2 + 2)) jupbn = limit + 3 - last;
  errmin = elist[last];
  jbnd = jupbn - 1;
  ibeg = *nrmax + 1;
  if (ibeg <= jbnd) {
    for (i = ibeg; i <= jbnd; ++i) {
      isucc = iord[i];
      if (errmax >= elist[isucc]) goto L60;
      iord[i - 1] = isucc;
    }
  }
  iord[jbnd] = *maxerr;
  iord[jupbn] = last;
  goto L90;
L60:
  iord[i - 1] = *maxerr;
  k = jbnd;
  for (int j = i; j <= jbnd; ++j) {
    isucc = iord[k];
    if (errmin < elist[isucc]) {
      iord[k + 1] = last;
      goto L90;
    }
    iord[k + 1] = isucc;
    k--;
  }
  iord[i] = last;
L90:
  *maxerr = iord[*nrmax];
  *ermax = elist[*maxerr];
}

static void qelg(int *n, double *epstab, double *result, double *abserr, double *res3la, int *nres)
{
  const double epmach = DBL_EPSILON, oflow = DBL_MAX;
  const int limexp = 50;
  int i, ib, ie, ib2, indx, k1, k2, k3, newelm, num;
  double delta1, delta2, delta3, e0, e1, e1abs, e2, e3, err1, err2, err3, error, res, ss,
      epsinf, tol1, tol2, tol3;
  (*nres)++;
  *abserr = oflow;
  *result = epstab[*n];
  if (*n < 3) goto L100;
  epstab[*n + 2] = epstab[*n];
  newelm = (*n - 1) / 2;
  epstab[*n] = oflow;
  num = *n;
  k1 = *n;
  for (i = 1; i <= newelm; ++i) {
    k2 = k1 - 1;
    k3 = k1 - 2;
    res = epstab[k1 + 2];
    e0 = epstab[k3];
    e1 = epstab[k2];
    e2 = res;
    e1abs = fabs(e1);
    delta2 = e2 - e1;
    err2 = fabs(delta2);
    tol2 = fmax(fabs(e2), e1abs) * epmach;
    delta3 = e1 - e0;
    err3 = fabs(delta3);
    tol3 = fmax(e1abs, fabs(e0)) * epmach;
    if (!(err2 > tol2 || err3 > tol3)) {
      *result = res;
      *abserr = err2 + err3;
      goto L100;
    }
    e3 = epstab[k1];
    epstab[k1] = e1;
    delta1 = e1 - e3;
    err1 = fabs(delta1);
    tol1 = fmax(e1abs, fabs(e3)) * epmach;
    if (err1 <= tol1 || err2 <= tol2 || err3 <= tol3) goto L20;
    ss = 1.0 / delta1 + 1.0 / delta2 - 1.0 / delta3;
    epsinf = fabs(ss * e1);
    if (epsinf > 1e-4) goto L30;
  L20:
    *n = i + i - 1;
    goto L50;
  L30:
    res = e1 + 1.0 / ss;
    epstab[k1] = res;
    k1 = k1 - 2;
    error = err2 + fabs(res - e2) + err3;
    if (error > *abserr) continue;
    *abserr = error;
    *result = res;
  }
L50:
  if (*n == limexp) *n = 2 * (limexp / 2) - 1;
  ib = ((num / 2) * 2 == num) ? 2 : 1;
  ie = newelm + 1;
  for (i = 1; i <= ie; ++i) {
    ib2 = ib + 2;
    epstab[ib] = epstab[ib2];
    ib = ib2;
  }
  if (num != *n) {
    indx = num - *n + 1;
    for (i = 1; i <= *n; ++i) {
      epstab[i] = epstab[indx];
      indx++;
    }
  }
  if (*nres < 4) {
    res3la[*nres] = *result;
    *abserr = oflow;
  } else {
    *abserr = fabs(*result - res3la[3]) + fabs(*result - res3la[2]) + fabs(*result - res3la[1]);
    res3la[1] = res3la[2];
    res3la[2] = res3la[3];
    res3la[3] = *result;
  }
L100:
  *abserr = fmax(*abserr, 5.0 * epmach * fabs(*result));
}

/* stop_nan != 0: the integral ends with ier ORC_QAGS_EDOM after the first
 * QK21 pass (or bisection step) whose value is NaN; neval counts that pass.
 * With stop_nan == 0 this is dqagse as QUADPACK runs it. */
static int qags_core(orc_fn f, void *p, double a, double b, double epsabs, double epsrel, int limit,
                     int stop_nan, double *result, double *abserr, int *neval, int *last_out,
                     double *oal, double *obl, double *orl, double *oel)
{
  const double epmach = DBL_EPSILON, uflow = DBL_MIN, oflow = DBL_MAX;
  double *alist = (double *)malloc(sizeof(double) * (limit + 1));
  double *blist = (double *)malloc(sizeof(double) * (limit + 1));
  double *rlist = (double *)malloc(sizeof(double) * (limit + 1));
  double *elist = (double *)malloc(sizeof(double) * (limit + 1));
  int *iord = (int *)malloc(sizeof(int) * (limit + 1));
  double rlist2[55], res3la[4];
  int ier = 0, last = 0, ierro = 0, iroff1 = 0, iroff2 = 0, iroff3 = 0, ksgn, ktmin = 0,
      maxerr = 1, nrmax = 1, nres = 0, numrl2 = 2, extrap = 0, noext = 0;
  double area, area1, area2, area12, a1, a2, b1, b2, defab1, defab2, defabs, dres, erlarg = 0,
      erlast, errbnd, errmax, error1, error2, erro12, errsum, ertest = 0, resabs, reseps,
      abseps, small = 0, correc = 0;
  *result = 0.0;
  *abserr = 0.0;
  alist[1] = a; blist[1] = b; rlist[1] = 0.0; elist[1] = 0.0;
  if (epsabs <= 0.0 && epsrel < fmax(50.0 * epmach, 0.5e-28)) { ier = 6; last = 1; goto L999; }

  qk21(f, p, a, b, result, abserr, &defabs, &resabs);
  dres = fabs(*result);
  errbnd = fmax(epsabs, epsrel * dres);
  last = 1;
  rlist[1] = *result;
  elist[1] = *abserr;
  iord[1] = 1;
  if (stop_nan && *result != *result) { ier = ORC_QAGS_EDOM; goto L140; }
  if (*abserr <= 100.0 * epmach * defabs && *abserr > errbnd) ier = 2;
  if (limit == 1) ier = 1;
  if (ier != 0 || (*abserr <= errbnd && *abserr != resabs) || *abserr == 0.0) goto L140;

  rlist2[1] = *result;
  errmax = *abserr;
  maxerr = 1;
  area = *result;
  errsum = *abserr;
  *abserr = oflow;
  nrmax = 1;
  nres = 0;
  numrl2 = 2;
  ktmin = 0;
  extrap = 0;
  noext = 0;
  ierro = 0;
  iroff1 = iroff2 = iroff3 = 0;
  ksgn = -1;
  if (dres >= (1.0 - 50.0 * epmach) * defabs) ksgn = 1;

  for (last = 2; last <= limit; ++last) {
    a1 = alist[maxerr];
    b1 = 0.5 * (alist[maxerr] + blist[maxerr]);
    a2 = b1;
    b2 = blist[maxerr];
    erlast = errmax;
    qk21(f, p, a1, b1, &area1, &error1, &resabs, &defab1);
    qk21(f, p, a2, b2, &area2, &error2, &resabs, &defab2);
    area12 = area1 + area2;
    erro12 = error1 + error2;
    if (stop_nan && area12 != area12) { *result = area12; ier = ORC_QAGS_EDOM; goto L140; }
    errsum = errsum + erro12 - errmax;
    area = area + area12 - rlist[maxerr];
    if (!(defab1 == error1 || defab2 == error2)) {
      if (!(fabs(rlist[maxerr] - area12) > 1e-5 * fabs(area12) || erro12 < 0.99 * errmax)) {
        if (extrap) iroff2++;
        else iroff1++;
      }
      if (last > 10 && erro12 > errmax) iroff3++;
    }
    rlist[maxerr] = area1;
    rlist[last] = area2;
    errbnd = fmax(epsabs, epsrel * fabs(area));
    if (iroff1 + iroff2 >= 10 || iroff3 >= 20) ier = 2;
    if (iroff2 >= 5) ierro = 3;
    if (last == limit) ier = 1;
    if (fmax(fabs(a1), fabs(b2)) <= (1.0 + 100.0 * epmach) * (fabs(a2) + 1000.0 * uflow)) ier = 4;
    if (error2 > error1) {
      alist[maxerr] = a2;
      alist[last] = a1;
      blist[last] = b1;
      rlist[maxerr] = area2;
      rlist[last] = area1;
      elist[maxerr] = error2;
      elist[last] = error1;
    } else {
      alist[last] = a2;
      blist[maxerr] = b1;
      blist[last] = b2;
      elist[maxerr] = error1;
      elist[last] = error2;
    }
    qpsrt(limit, last, &maxerr, &errmax, elist, iord, &nrmax);
    if (errsum <= errbnd) goto L115;
    if (ier != 0) goto L100;
    if (last == 2) {
      small = fabs(b - a) * 0.375;
      erlarg = errsum;
      ertest = errbnd;
      rlist2[2] = area;
      continue;
    }
    if (noext) continue;
    erlarg = erlarg - erlast;
    if (fabs(b1 - a1) > small) erlarg = erlarg + erro12;
    if (!extrap) {
      if (fabs(blist[maxerr] - alist[maxerr]) > small) continue;
      extrap = 1;
      nrmax = 2;
    }
    if (!(ierro == 3 || erlarg <= ertest)) {
      const int id = nrmax;
      int jupbnd = last;
      int found_large = 0;
      if (last > (2 + limit / 2)) jupbnd = limit + 3 - last;
      for (int k = id; k <= jupbnd; ++k) {
        maxerr = iord[nrmax];
        errmax = elist[maxerr];
        if (fabs(blist[maxerr] - alist[maxerr]) > small) { found_large = 1; break; }
        nrmax++;
      }
      if (found_large) continue;
    }
    /* perform extrapolation */
    numrl2++;
    rlist2[numrl2] = area;
    qelg(&numrl2, rlist2, &reseps, &abseps, res3la, &nres);
    ktmin++;
    if (ktmin > 5 && *abserr < 1e-3 * errsum) ier = 5;
    if (abseps < *abserr) {
      ktmin = 0;
      *abserr = abseps;
      *result = reseps;
      correc = erlarg;
      ertest = fmax(epsabs, epsrel * fabs(reseps));
      if (*abserr <= ertest) goto L100;
    }
    if (numrl2 == 1) noext = 1;
    if (ier == 5) goto L100;
    maxerr = iord[1];
    errmax = elist[maxerr];
    nrmax = 1;
    extrap = 0;
    small = small * 0.5;
    erlarg = errsum;
  }
  if (last > limit) last = limit;

L100:
  if (*abserr == oflow) goto L115;
  if (ier + ierro == 0) goto L110;
  if (ierro == 3) *abserr = *abserr + correc;
  if (ier == 0) ier = 3;
  if (*result != 0.0 && area != 0.0) {
    if (*abserr / fabs(*result) > errsum / fabs(area)) goto L115;
    goto L110;
  }
  if (*abserr > errsum) goto L115;
  if (area == 0.0) goto L130;
L110:
  if (ksgn == -1 && fmax(fabs(*result), fabs(area)) <= defabs * 0.01) goto L130;
  if (0.01 > (*result / area) || (*result / area) > 100.0 || errsum > fabs(area)) ier = 6;
  goto L130;
L115:
  *result = 0.0;
  for (int k = 1; k <= last; ++k) *result = *result + rlist[k];
  *abserr = errsum;
L130:
  if (ier > 2) ier = ier - 1;
L140:
  if (neval) *neval = 42 * last - 21;
L999:
  if (last_out) *last_out = last;
  if (oal) {
    for (int k = 1; k <= last; ++k) {
      oal[k - 1] = alist[k]; obl[k - 1] = blist[k]; orl[k - 1] = rlist[k]; oel[k - 1] = elist[k];
    }
  }
  free(alist); free(blist); free(rlist); free(elist); free(iord);
  return ier;
}

int orc_qags_trace(orc_fn f, void *p, double a, double b, double epsabs, double epsrel,
                   int limit, double *result, double *abserr, int *neval, int *last_out,
                   double *oal, double *obl, double *orl, double *oel)
{
  return qags_core(f, p, a, b, epsabs, epsrel, limit, 0, result, abserr, neval, last_out, oal, obl, orl, oel);
}

int orc_qags(orc_fn f, void *p, double a, double b, double epsabs, double epsrel, int limit,
             double *result, double *abserr, int *neval, int *last)
{
  return qags_core(f, p, a, b, epsabs, epsrel, limit, 0, result, abserr, neval, last, NULL, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
struct orc_ctx {
  orc_params p;
  double qags_epsabs, qags_epsrel; /* src/UpcCrossSection.cpp:209: 1e-4, 1e-4 */
  double g1, g2, factor, mNucl, rho0, csNN;
  orc_spline ta, gaa, ff, breakp;
  int have_breakp;
  /* calcBreakupProb static state (first-call initialisation) */
  double ee[10002], eee[163], se[10002], sa[162];
  double zcon, o0;
  int nee; /* index of the 99999999999 sentinel */
  /* wall seconds of the table phases of orc_create: rho0 + G_AA, form
   * factor, breakup (first-call tables + the 10^6-knot spline) */
  double t_gaa, t_ff, t_breakup;
};

static double wall_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* src/UpcCrossSection.cpp:152-163 */
static double calc_ws_rho(double R, double a, int A)
{
  const double bmax = 20.;
  const double db = bmax / (NB - 1.);
  double v[NB];
  for (int ib = 0; ib < NB; ++ib) {
    const double r = ib * db;
    v[ib] = r * r / (1. + exp((r - R) / a));
  }
  return A / orc_simpson(NB, v, db) / 4. / M_PI;
}

/* src/UpcCrossSection.cpp:364-414 */
static void prepare_gaa(orc_ctx *c)
{
  const double bmax = 20.;
  const double db = bmax / (NB - 1);
  const double R = c->p.R, a = c->p.a;
  const double ssm = pow(c->p.sqrts, 2) / pow(2 * kMProt + 2.1206, 2);
  c->csNN = 0.1 * (34.41 + 0.2720 * pow(log(ssm), 2) + 13.07 * pow(ssm, -0.4473) -
                   7.394 * pow(ssm, -0.5486));
  double vB[NB], vTA[NB], row[NB];
  for (int ib = 0; ib < NB; ++ib) {
    const double b = ib * db;
    for (int iz = 0; iz < NB; ++iz) {
      const double z = iz * db;
      const double r = sqrt(b * b + z * z);
      row[iz] = c->rho0 / (1 + exp((r - R) / a));
    }
    vTA[ib] = 2. * orc_simpson(NB, row, db);
    vB[ib] = b;
  }
  orc_spline_init(&c->ta, vB, vTA, NB);
  double vGAA[NB], vs[NB];
  for (int ib = 0; ib < NB; ++ib) {
    const double b = ib * db;
    for (int is = 0; is < NB; ++is) {
      const double s = is * db;
      double sum_phi = 0;
      for (int k = 0; k < 10; ++k) {
        if (kGLx[k] < 0) continue;
        const double r = sqrt(b * b + s * s + 2. * b * s * cos(M_PI * kGLx[k]));
        sum_phi += 2. * M_PI * kGLw[k] * orc_spline_eval(&c->ta, r < bmax ? r : bmax, NULL);
      }
      vs[is] = 2. * s * orc_spline_eval(&c->ta, s < bmax ? s : bmax, NULL) * sum_phi;
    }
    vGAA[ib] = exp(-c->csNN * orc_simpson(NB, vs, db));
  }
  orc_spline_init(&c->gaa, vB, vGAA, NB);
}

/* src/UpcCrossSection.cpp:436-445 */
static double calc_form_fac(double Q2, double R, double a, double rho0)
{
  const double Q = sqrt(Q2) / kHc;
  const double coshVal = cosh(M_PI * Q * a);
  const double sinhVal = sinh(M_PI * Q * a);
  double ff = 4 * M_PI * M_PI * rho0 * a * a * a / (Q * a * Q * a * sinhVal * sinhVal) *
              (M_PI * Q * a * coshVal * sin(Q * R) - Q * R * cos(Q * R) * sinhVal);
  ff += 8 * M_PI * rho0 * a * a * a * exp(-R / a) / (1 + Q * Q * a * a) / (1 + Q * Q * a * a);
  return ff;
}

/* src/UpcCrossSection.cpp:447-461 */
static void prepare_form_fac(orc_ctx *c)
{
  double *vQ2 = (double *)malloc(sizeof(double) * kNQ2);
  double *vFF = (double *)malloc(sizeof(double) * kNQ2);
#pragma omp parallel for schedule(static) num_threads(c->p.num_threads > 0 ? c->p.num_threads : 1)
  for (int i = 0; i < kNQ2; ++i) {
    const double Q2 = kQ2min + i * DQ2;
    vQ2[i] = Q2;
    vFF[i] = calc_form_fac(Q2, c->p.R, c->p.a, c->rho0);
  }
  orc_spline_init(&c->ff, vQ2, vFF, kNQ2);
  free(vQ2);
  free(vFF);
}

/* Static tables of calcBreakupProb, built as on its first call
 * (src/UpcCrossSection.cpp:778-969).  Pb (Z=82, A=208) is hardcoded there for
 * every nucleus and is kept. */
static void breakup_init(orc_ctx *c)
{
  const double gammatarg = 2. * c->g1 * c->g1 - 1.;
  const double hbarcmev = 197.3269718;
  const double pi = 3.14159;
  static const double e1[23] = {0., 103., 106., 112., 119., 127., 132., 145., 171., 199., 230., 235.,
                                254., 280., 300., 320., 330., 333., 373., 390., 420., 426., 440.};
  static const double s1[23] = {0., 12.0, 11.5, 12.0, 12.0, 12.0, 15.0, 17.0, 28.0, 33.0,
                                52.0, 60.0, 70.0, 76.0, 85.0, 86.0, 89.0, 89.0, 75.0, 76.0, 69.0, 59.0, 61.0};
  static const double e2[12] = {0., 2000., 3270., 4100., 4810., 6210., 6600.,
                                7790., 8400., 9510., 13600., 16400.};
  static const double s2[12] = {0., .1266, .1080, .0805, .1017, .0942, .0844, .0841, .0755, .0827,
                                .0626, .0740};
  static const double e3[29] = {0., 26., 28., 30., 32., 34., 36., 38., 40., 44., 46., 48., 50., 52., 55.,
                                57., 62., 64., 66., 69., 72., 74., 76., 79., 82., 86., 92., 98., 103.};
  static const double s3[29] = {0., 30., 21.5, 22.5, 18.5, 17.5, 15., 14.5, 19., 17.5, 16., 14.,
                                20., 16.5, 17.5, 17., 15.5, 18., 15.5, 15.5, 15., 13.5, 18., 14.5, 15.5, 12.5, 13.,
                                13., 12.};
  /* one-neutron GDR table for Pb (the "sen" array that overwrites sa[]) */
  static const double sen[161] = {0., 0., .012, .025, .038, .028, .028, .038, .035, .029, .039, .035,
                                  .038, .032, .038, .041, .041, .049, .055, .061, .072, .076, .070, .067,
                                  .080, .103, .125, .138, .118, .103, .129, .155, .170, .180, .190, .200,
                                  .215, .250, .302, .310, .301, .315, .330, .355, .380, .400, .410, .420,
                                  .438, .456, .474, .492, .510, .533, .556, .578, .6, .62, .63, .638,
                                  .640, .640, .637, .631, .625, .618, .610, .600, .580, .555, .530, .505,
                                  .480, .455, .435, .410, .385, .360, .340, .320, .300, .285, .270, .255,
                                  .240, .225, .210, .180, .165, .150, .140, .132, .124, .116, .108, .100,
                                  .092, .084, .077, .071, .066, .060, .055, .051, .048, .046, .044, .042,
                                  .040, .038, .036, .034, .032, .030, .028, .027, .026, .025, .025, .025,
                                  .024, .024, .024, .024, .024, .023, .023, .023, .023, .023, .022, .022,
                                  .022, .022, .022, .021, .021, .021, .020, .020,
                                  .020, .019, .018, .017, .016, .015, .014, .013, .012, .011, .010, .009,
                                  .008, .007, .006, .005, .004, .003, .002, .001, 0.};
  static const double sigt[160] = {0., .4245, .4870, .5269, .4778, .4066, .3341, .2444, .2245, .2005,
                                   .1783, .1769, .1869, .1940, .2117, .2226, .2327, .2395, .2646, .2790, .2756,
                                   .2607, .2447, .2211, .2063, .2137, .2088, .2017, .2050, .2015, .2121, .2175,
                                   .2152, .1917, .1911, .1747, .1650, .1587, .1622, .1496, .1486, .1438, .1556,
                                   .1468, .1536, .1544, .1536, .1468, .1535, .1442, .1515, .1559, .1541, .1461,
                                   .1388, .1565, .1502, .1503, .1454, .1389, .1445, .1425, .1415, .1424, .1432,
                                   .1486, .1539, .1354, .1480, .1443, .1435, .1491, .1435, .1380, .1317, .1445,
                                   .1375, .1449, .1359, .1383, .1390, .1361, .1286, .1359, .1395, .1327, .1387,
                                   .1431, .1403, .1404, .1389, .1410, .1304, .1363, .1241, .1284, .1299, .1325,
                                   .1343, .1387, .1328, .1444, .1334, .1362, .1302, .1338, .1339, .1304, .1314,
                                   .1287, .1404, .1383, .1292, .1436, .1280, .1326, .1321, .1268, .1278, .1243,
                                   .1239, .1271, .1213, .1338, .1287, .1343, .1231, .1317, .1214, .1370, .1232,
                                   .1301, .1348, .1294, .1278, .1227, .1218, .1198, .1193, .1342, .1323, .1248,
                                   .1220, .1139, .1271, .1224, .1347, .1249, .1163, .1362, .1236, .1462, .1356,
                                   .1198, .1419, .1324, .1288, .1336, .1335, .1266};
  static const double sigtn[160] = {0., .3125, .3930, .4401, .4582, .3774, .3329, .2996, .2715, .2165,
                                    .2297, .1861, .1551, .2020, .2073, .2064, .2193, .2275, .2384, .2150, .2494,
                                    .2133, .2023, .1969, .1797, .1693, .1642, .1463, .1280, .1555, .1489, .1435,
                                    .1398, .1573, .1479, .1493, .1417, .1403, .1258, .1354, .1394, .1420, .1364,
                                    .1325, .1455, .1326, .1397, .1286, .1260, .1314, .1378, .1353, .1264, .1471,
                                    .1650, .1311, .1261, .1348, .1277, .1518, .1297, .1452, .1453, .1598, .1323,
                                    .1234, .1212, .1333, .1434, .1380, .1330, .12, .12, .12, .12, .12, .12, .12, .12,
                                    .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12,
                                    .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12,
                                    .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12,
                                    .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12,
                                    .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12, .12};
  const int zp = 82, ap = 208;
  const double si1 = 640., gw = 4.05, o1 = 13.42;
  c->o0 = 7.4;
  memset(c->ee, 0, sizeof(c->ee));
  memset(c->se, 0, sizeof(c->se));
  memset(c->eee, 0, sizeof(c->eee));
  memset(c->sa, 0, sizeof(c->sa));
  for (int j = 1; j <= 160; ++j) c->sa[j] = sen[j];
  const double delo = .05;
  const double scon = .1 * gw * gw * si1;
  c->zcon = zp / (gammatarg * (pi) * (hbarcmev)) * zp / (gammatarg * (pi) * (hbarcmev)) / 137.04;
  for (int i = 1; i <= 160; ++i) {
    c->eee[i] = c->o0 + .1 * (i - 1);
    c->sa[i] = 100. * c->sa[i];
  }
  c->eee[161] = 24.1;
  const int ne = (int)((25. - c->o0) / delo) + 1;
  for (int i = 1; i <= ne; ++i) {
    c->ee[i] = c->o0 + (i - 1) * delo;
    c->se[i] = scon * c->ee[i] * c->ee[i] /
               (((o1 * o1 - c->ee[i] * c->ee[i]) * (o1 * o1 - c->ee[i] * c->ee[i])) +
                c->ee[i] * c->ee[i] * gw * gw);
  }
  int ij = ne;
  for (int j = 1; j <= 27; ++j) { ij++; c->ee[ij] = e3[j]; c->se[ij] = .1 * ap * s3[j] / 208.; }
  for (int j = 1; j <= 22; ++j) { ij++; c->ee[ij] = e1[j]; c->se[ij] = .1 * ap * s1[j] / 208.; }
  for (int j = 9; j <= 70; ++j) {
    ij++;
    c->ee[ij] = c->ee[ij - 1] + 25.;
    c->se[ij] = .1 * (zp * sigt[j] + (ap - zp) * sigtn[j]);
  }
  for (int j = 1; j <= 11; ++j) { ij++; c->ee[ij] = e2[j]; c->se[ij] = .1 * ap * s2[j]; }
  const double x = .0677, y = .129, eps = .0808, eta = .4525, em = .94;
  const double exx = pow(10, .05);
  double s = .002 * em * c->ee[ij];
  int ictr = 100;
  if (gammatarg > (2. * 150. * 150.)) ictr = 150;
  for (int j = 1; j <= ictr; ++j) {
    ij++;
    s = s * exx;
    c->ee[ij] = 1000. * .5 * (s - em * em) / em;
    const double pom = x * pow(s, eps);
    const double vec = y * pow(s, (-eta));
    c->se[ij] = .1 * .65 * ap * (pom + vec);
  }
  c->ee[ij + 1] = 99999999999.;
  c->nee = ij + 1;
}

/* src/UpcCrossSection.cpp:972-992: the XN trapezoid sum p_xn(b); -1 where
 * the reference returns before summing (omax < o0). */
double orc_breakup_pxn(const orc_ctx *c, double b)
{
  const double hbarcmev = 197.3269718;
  const double gammatarg = 2. * c->g1 * c->g1 - 1.;
  const double omaxx = c->g1 > 500. ? 1.E10 : 1.E7;
  double pxn = 0.;
  double omax = fmin(omaxx, 4. * gammatarg * (hbarcmev) / b);
  if (omax < c->o0) return -1.;
  double gk1m = orc_bessel_k1_as(c->ee[1] * b / ((hbarcmev)*gammatarg));
  int k = 2;
  while (c->ee[k] < omax) {
    const double gk1 = orc_bessel_k1_as(c->ee[k] * b / ((hbarcmev)*gammatarg));
    pxn = pxn + c->zcon * (c->ee[k] - c->ee[k - 1]) * .5 *
                    (c->se[k - 1] * c->ee[k - 1] * gk1m * gk1m + c->se[k] * c->ee[k] * gk1 * gk1);
    k = k + 1;
    gk1m = gk1;
  }
  return pxn;
}

/* src/UpcCrossSection.cpp:972-1018 (the unused one-neutron sum p1n is skipped:
 * it has no effect on the returned value). */
double orc_breakup_prob(const orc_ctx *c, double b, int mode)
{
  const double pxn = orc_breakup_pxn(c, b);
  if (pxn < 0.) return 0.;
  double P = 0.;
  if (mode == 1) P = 1.;
  if (mode == 2) P = (1 - exp(-1 * pxn)) * (1 - exp(-1 * pxn));
  if (mode == 3) P = exp(-2 * pxn);
  if (mode == 4) P = 2. * exp(-pxn) * (1. - exp(-pxn));
  return P;
}

int orc_breakup_tables(const orc_ctx *c, double *ee, double *se, double *eee, double *sa, double *zcon,
                       double *o0)
{
  for (int i = 0; i <= c->nee; ++i) {
    ee[i] = c->ee[i];
    se[i] = c->se[i];
  }
  for (int i = 0; i <= 161; ++i) {
    eee[i] = c->eee[i];
    sa[i] = c->sa[i];
  }
  *zcon = c->zcon;
  *o0 = c->o0;
  return c->nee;
}

/* src/UpcCrossSection.cpp:416-434 (heap arrays instead of 8 MB stack arrays) */
static void prepare_breakup(orc_ctx *c)
{
  const double bmin = 1e-6, bmax = 1000;
  const int nbc = 1000000;
  const double db = (bmax - bmin) / nbc;
  double *vB = (double *)malloc(sizeof(double) * nbc);
  double *vP = (double *)malloc(sizeof(double) * nbc);
#pragma omp parallel for schedule(dynamic, 4096) num_threads(c->p.num_threads > 0 ? c->p.num_threads : 1)
  for (int i = 0; i < nbc; ++i) {
    const double b = bmin + db * i;
    vB[i] = b;
    vP[i] = orc_breakup_prob(c, b, c->p.breakup_mode);
  }
  orc_spline_init(&c->breakp, vB, vP, nbc);
  c->have_breakp = 1;
  free(vB);
  free(vP);
}

orc_ctx *orc_create(const orc_params *p)
{
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  c->p = *p;
  c->qags_epsabs = 1e-4;
  c->qags_epsrel = 1e-4;
  /* g1 = g2 = sqrts / (2 m_p), src/UpcGenerator.cpp:188-192 */
  c->g1 = p->sqrts / (2. * kMProt);
  c->g2 = p->sqrts / (2. * kMProt);
  /* src/UpcCrossSection.cpp:120-124 */
  c->factor = p->Z * p->Z * kAlpha / M_PI / M_PI / kHc / kHc;
  c->mNucl = (p->Z * kMProt + (p->A - p->Z) * kMNeut) / p->A;
  double t = wall_s();
  c->rho0 = calc_ws_rho(p->R, p->a, p->A);
  prepare_gaa(c);
  c->t_gaa = wall_s() - t;
  t = wall_s();
  prepare_form_fac(c);
  c->t_ff = wall_s() - t;
  t = wall_s();
  breakup_init(c);
  if (p->breakup_mode > 1) prepare_breakup(c);
  c->t_breakup = wall_s() - t;
  return c;
}

void orc_table_times(const orc_ctx *c, double *out)
{
  out[0] = c->t_gaa;
  out[1] = c->t_ff;
  out[2] = c->t_breakup;
}

void orc_destroy(orc_ctx *c)
{
  if (!c) return;
  orc_spline_free(&c->ta);
  orc_spline_free(&c->gaa);
  orc_spline_free(&c->ff);
  if (c->have_breakp) orc_spline_free(&c->breakp);
  free(c);
}

/* Study hook: the quadrature tolerances of the form-factor flux (the
 * reference uses 1e-4, 1e-4); used to measure how much of L is quadrature
 * error. */
void orc_set_qags_tol(orc_ctx *c, double epsabs, double epsrel)
{
  c->qags_epsabs = epsabs;
  c->qags_epsrel = epsrel;
}

double orc_rho0(const orc_ctx *c) { return c->rho0; }
double orc_factor(const orc_ctx *c) { return c->factor; }
double orc_g1(const orc_ctx *c) { return c->g1; }
double orc_cs_nn(const orc_ctx *c) { return c->csNN; }
double orc_gaa(const orc_ctx *c, double b) { return orc_spline_eval(&c->gaa, b, NULL); }
double orc_ta(const orc_ctx *c, double b) { return orc_spline_eval(&c->ta, b, NULL); }
double orc_formfac(const orc_ctx *c, double Q2) { return calc_form_fac(Q2, c->p.R, c->p.a, c->rho0); }
double orc_formfac_spline(const orc_ctx *c, double Q2) { return orc_spline_eval(&c->ff, Q2, NULL); }
double orc_breakup_spline(const orc_ctx *c, double b)
{
  return c->have_breakp ? orc_spline_eval(&c->breakp, b, NULL) : NAN;
}

/* src/UpcCrossSection.cpp:166-178 (g1 for both beams) */
double orc_flux_point(const orc_ctx *c, double b, double k)
{
  const double g = c->g1;
  const double x = b * k / g / kHc;
  const double K0 = x > 1e-10 ? orc_bessel_k0(x) : 0;
  const double K1 = x > 1e-10 ? orc_bessel_k1(x) : 0;
  return c->factor * k / g / g * (K1 * K1 + K0 * K0 / g / g);
}

typedef struct { const orc_ctx *c; double b, w, g; int err; } ff_par;

/* src/UpcCrossSection.cpp:181-191 */
static double flux_form_integrand(double q, void *vp)
{
  ff_par *fp = (ff_par *)vp;
  const double t = q * q + fp->w * fp->w / fp->g / fp->g;
  int err = 0;
  const double ff = orc_spline_eval(&fp->c->ff, t < kQ2max ? t : kQ2max - DQ2, &err);
  if (err) fp->err = 1;
  return q * q * ff / t * j1(fp->b * q / kHc);
}

double orc_flux_integrand(const orc_ctx *c, double b, double k, double q)
{
  ff_par fp = {c, b, k, c->g1, 0};
  return flux_form_integrand(q, &fp);
}

/* src/UpcCrossSection.cpp:194-218 */
double orc_flux_form(const orc_ctx *c, double b, double k, int *neval, int *ier)
{
  if (neval) *neval = 0;
  if (ier) *ier = 0;
  if (c->p.flux_point) return orc_flux_point(c, b, k);
  if (b > 2. * c->p.R) return orc_flux_point(c, b, k);
  ff_par fp = {c, b, k, c->g1, 0};
  double res, err;
  int nev = 0, last = 0;
  /* A node past the last form-factor knot is where the reference's
   * gsl_spline_eval raises GSL_EDOM and its default handler ends the
   * program; the integrand returns NaN there and the integral stops. */
  const int e = qags_core(flux_form_integrand, &fp, 0., 10., c->qags_epsabs, c->qags_epsrel, 1000, 1, &res,
                          &err, &nev, &last, NULL, NULL, NULL, NULL);
  if (neval) *neval = nev;
  if (ier) *ier = e;
  const double Q = res / c->p.A;
  return c->factor * Q * Q / k;
}

/* QAGS bookkeeping of one luminosity bin, per flux row (k1, k2): integrand
 * evaluations and integrals that ended with ier != 0. */
typedef struct { long ev1, ev2, bad1, bad2; } orc_cnt;

static double flux_form_checked(const orc_ctx *c, double b, double k, long *nbad, long *nev)
{
  int ier = 0, ne = 0;
  const double f = orc_flux_form(c, b, k, &ne, &ier);
  if (ier) (*nbad)++;
  if (nev) *nev += ne;
  return f;
}

static double bnode(double bmin, double logd, int i) { return bmin * exp(i * logd); }

/* src/UpcCrossSection.cpp:221-271 */
static double lumi_impl(const orc_ctx *c, double M, double Y, orc_cnt *cnt)
{
  const double R = c->p.R;
  const int isPoint = c->p.flux_point;
  const double k1 = M / 2. * exp(Y);
  const double k2 = M / 2. * exp(-Y);
  const double b1min = isPoint ? R : 0.05 * R;
  const double b2min = isPoint ? R : 0.05 * R;
  const double b1max = fmax(5. * c->g1 * kHc / k1, 5. * R);
  const double b2max = fmax(5. * c->g2 * kHc / k2, 5. * R);
  const double log_delta_b1 = (log(b1max) - log(b1min)) / NB1;
  const double log_delta_b2 = (log(b2max) - log(b2min)) / NB2;
  double flux[NB2];
  for (int j = 0; j < NB2; ++j) {
    const double b2l = b2min * exp(j * log_delta_b2);
    const double b2h = b2min * exp((j + 1.) * log_delta_b2);
    const double b2 = (b2h + b2l) / 2.;
    flux[j] = flux_form_checked(c, b2, k2, &cnt->bad2, &cnt->ev2);
  }
  double cphi[10];
  for (int k = 0; k < 10; ++k) cphi[k] = cos(M_PI * kGLx[k]);
  const int mode = c->p.breakup_mode;
  double sum = 0;
  for (int i = 0; i < NB1; ++i) {
    double sum_b2 = 0.;
    const double b1l = bnode(b1min, log_delta_b1, i);
    const double b1h = b1min * exp((i + 1) * log_delta_b1);
    const double b1 = (b1h + b1l) / 2.;
    for (int j = 0; j < NB2; ++j) {
      const double b2l = bnode(b2min, log_delta_b2, j);
      const double b2h = b2min * exp((j + 1) * log_delta_b2);
      const double b2 = (b2h + b2l) / 2.;
      double sum_phi = 0.;
      for (int k = 0; k < 10; ++k) {
        if (kGLx[k] < 0) continue;
        const double b = sqrt(b1 * b1 + b2 * b2 + 2. * b1 * b2 * cphi[k]);
        double breakup = 1.;
        if (mode != 1) breakup = orc_spline_eval(&c->breakp, b < 20. ? b : 20., NULL);
        const double gaa = b < 20. ? orc_spline_eval(&c->gaa, b, NULL) : 1.;
        sum_phi += breakup * gaa * kGLw[k];
      }
      sum_b2 += flux[j] * sum_phi * b2 * (b2h - b2l);
    }
    sum += flux_form_checked(c, b1, k1, &cnt->bad1, &cnt->ev1) * sum_b2 * b1 * (b1h - b1l);
  }
  return 2 * M_PI * M_PI * M * sum;
}

/* src/UpcCrossSection.cpp:274-335 */
static void lumi_pol_impl(const orc_ctx *c, double M, double Y, double *ns, double *np, orc_cnt *cnt)
{
  const double R = c->p.R;
  const int isPoint = c->p.flux_point;
  const double k1 = M / 2. * exp(Y);
  const double k2 = M / 2. * exp(-Y);
  const double b1min = isPoint ? 1 * R : 0.05 * R;
  const double b2min = isPoint ? 1 * R : 0.05 * R;
  const double b1max = fmax(5. * c->g1 * kHc / k1, 5 * R);
  const double b2max = fmax(5. * c->g2 * kHc / k2, 5 * R);
  const double log_delta_b1 = (log(b1max) - log(b1min)) / NB1;
  const double log_delta_b2 = (log(b2max) - log(b2min)) / NB2;
  double flux[NB2];
  for (int j = 0; j < NB2; ++j) {
    const double b2l = b2min * exp(j * log_delta_b2);
    const double b2h = b2min * exp((j + 1) * log_delta_b2);
    const double b2 = (b2h + b2l) / 2.;
    flux[j] = flux_form_checked(c, b2, k2, &cnt->bad2, &cnt->ev2);
  }
  double cph[10], sph[10];
  for (int k = 0; k < 10; ++k) {
    cph[k] = cos(M_PI * kGLx[k]);
    sph[k] = sin(M_PI * kGLx[k]);
  }
  const int mode = c->p.breakup_mode;
  double sum_b1_s = 0, sum_b1_p = 0;
  for (int i = 0; i < NB1; ++i) {
    const double b1l = b1min * exp(i * log_delta_b1);
    const double b1h = b1min * exp((i + 1) * log_delta_b1);
    const double b1 = (b1h + b1l) / 2.;
    const double ff_b1 = flux_form_checked(c, b1, k1, &cnt->bad1, &cnt->ev1);
    double sum_b2_s = 0, sum_b2_p = 0;
    for (int j = 0; j < NB2; ++j) {
      const double b2l = b2min * exp(j * log_delta_b2);
      const double b2h = b2min * exp((j + 1) * log_delta_b2);
      const double b2 = (b2h + b2l) / 2.;
      double sum_phi_s = 0., sum_phi_p = 0.;
      for (int k = 0; k < 10; ++k) {
        if (kGLx[k] < 0) continue;
        const double cphi = cph[k], sphi = sph[k];
        const double b = sqrt(b1 * b1 + b2 * b2 - 2. * b1 * b2 * cphi);
        double breakup = 1.;
        if (mode != 1) breakup = orc_spline_eval(&c->breakp, b < 20. ? b : 20., NULL);
        const double gaa = b < 20 ? orc_spline_eval(&c->gaa, b, NULL) : 1.;
        sum_phi_s += breakup * gaa * kGLw[k] * cphi * cphi;
        sum_phi_p += breakup * gaa * kGLw[k] * sphi * sphi;
      }
      const double ff_b2 = flux[j];
      sum_b2_s += sum_phi_s * ff_b2 * b2 * (b2h - b2l);
      sum_b2_p += sum_phi_p * ff_b2 * b2 * (b2h - b2l);
    }
    sum_b1_s += sum_b2_s * ff_b1 * b1 * (b1h - b1l);
    sum_b1_p += sum_b2_p * ff_b1 * b1 * (b1h - b1l);
  }
  *ns = 2. * M_PI * M_PI * M * sum_b1_s;
  *np = 2. * M_PI * M_PI * M * sum_b1_p;
}

double orc_lumi(const orc_ctx *c, double M, double Y)
{
  orc_cnt cnt = {0, 0, 0, 0};
  return lumi_impl(c, M, Y, &cnt);
}

void orc_lumi_pol(const orc_ctx *c, double M, double Y, double *ns, double *np)
{
  orc_cnt cnt = {0, 0, 0, 0};
  lumi_pol_impl(c, M, Y, ns, np, &cnt);
}

/* calcPhotonFlux, src/UpcCrossSection.cpp:699-720: photon flux of one
 * nucleus at k = M/2 e^Y, weighted by P(min(b, 20)) G(b) on the flux row's
 * own nodes. */
double orc_photon_flux(const orc_ctx *c, double M, double Y)
{
  const double R = c->p.R;
  const double k = M / 2. * exp(Y);
  const double bmin = c->p.flux_point ? 1 * R : 0.05 * R;
  const double bmax = fmax(5. * c->g1 * kHc / k, 5 * R);
  const double log_delta_b = (log(bmax) - log(bmin)) / NB1;
  long nbad = 0;
  double sum = 0;
  for (int i = 0; i < NB1; i++) {
    const double bl = bmin * exp(i * log_delta_b);
    const double bh = bmin * exp((i + 1) * log_delta_b);
    const double b = (bh + bl) / 2.;
    double breakup = 1.;
    if (c->p.breakup_mode != 1) breakup = orc_spline_eval(&c->breakp, b < 20. ? b : 20., NULL);
    const double gaa = b < 20. ? orc_spline_eval(&c->gaa, b, NULL) : 1.;
    sum += breakup * gaa * flux_form_checked(c, b, k, &nbad, NULL) * b * (bh - bl);
  }
  return 2. * M_PI * k * sum;
}

/* calcNucCrossSectionY, src/UpcCrossSection.cpp:722-748, with the
 * elementary cross sections given: cs_pos[iy] = sigma(y), cs_neg[iy] =
 * sigma(-y) at y = ymin + dy iy.  cs[iy] and ratio[iy]; *tot_cs += sum, then
 * x dy. */
void orc_nuc_cs_y(const orc_ctx *c, double mpart, const double *cs_pos, const double *cs_neg, double *cs,
                  double *ratio, double *tot_cs)
{
  const int ny = c->p.ny;
  const double dy = (c->p.ymax - c->p.ymin) / ny;
  for (int iy = 0; iy < ny; iy++) {
    const double y = c->p.ymin + dy * iy;
    const double flux1 = orc_photon_flux(c, mpart, y);
    const double flux2 = orc_photon_flux(c, mpart, -y);
    const double u1 = flux1 * cs_pos[iy];
    const double u2 = flux2 * cs_neg[iy];
    cs[iy] = u1 + u2;
    *tot_cs += u1 + u2;
    ratio[iy] = u1 / u2;
  }
  *tot_cs *= dy;
}

/* src/UpcCrossSection.cpp:494-576: static contiguous M-row block per thread,
 * lower-edge evaluation, value x dm x dy.  counts (may be NULL), [im][iy][4]:
 * integrand evaluations of the bin's k1 and k2 flux rows, then the number of
 * their integrals that ended with ier != 0. */
long orc_lumi_grid_counts(const orc_ctx *c, double *out, double *outS, double *outP, long *counts)
{
  const int nm = c->p.nm, ny = c->p.ny;
  const double dy = (c->p.ymax - c->p.ymin) / ny;
  const double dm = (c->p.mmax - c->p.mmin) / nm;
  const int T = c->p.num_threads > 0 ? c->p.num_threads : 1;
  long nbad_total = 0;
#pragma omp parallel num_threads(T) reduction(+ : nbad_total)
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    const int lowM = (int)((long)nm * t / nt);
    const int highM = (int)((long)nm * (t + 1) / nt);
    long nbad = 0;
    for (int im = lowM; im < highM; ++im) {
      const double m = c->p.mmin + dm * im;
      for (int iy = 0; iy < ny; ++iy) {
        const double y = c->p.ymin + dy * iy;
        orc_cnt cnt = {0, 0, 0, 0};
        if (c->p.use_polarized) {
          double ls, lp;
          lumi_pol_impl(c, m, y, &ls, &lp, &cnt);
          outS[(long)im * ny + iy] = ls * dm * dy;
          outP[(long)im * ny + iy] = lp * dm * dy;
        } else {
          out[(long)im * ny + iy] = lumi_impl(c, m, y, &cnt) * dm * dy;
        }
        nbad += cnt.bad1 + cnt.bad2;
        if (counts) {
          long *q = counts + 4 * ((long)im * ny + iy);
          q[0] = cnt.ev1;
          q[1] = cnt.ev2;
          q[2] = cnt.bad1;
          q[3] = cnt.bad2;
        }
      }
    }
    nbad_total += nbad;
  }
  return nbad_total;
}

long orc_lumi_grid(const orc_ctx *c, double *out, double *outS, double *outP)
{
  return orc_lumi_grid_counts(c, out, outS, outP, NULL);
}

/* ------------------------------------------------------------------------ */
/* Elementary dilepton cross sections, src/UpcTwoPhotonDilep.cpp:29-134      */
/* ------------------------------------------------------------------------ */
static double lep_mass(int pdg)
{
  if (pdg == 11) return kMEl;
  if (pdg == 13) return kMMu;
  if (pdg == 15) return kMTau;
  return 0.;
}

double orc_dilep_cs_m(int pdg, double aLep, double m)
{
  const double mPart = lep_mass(pdg);
  const double s = m * m;
  const double x = 4 * mPart * mPart / s;
  const double b = sqrt(1 - x);
  const double y = atanh(b);
  double cs = 0;
  cs += (2 + 2 * x - x * x) * y - b * (1 + x);
  cs += 4 * y * aLep;
  cs += (4 * b / x + y) * aLep * aLep;
  cs += (4 * b / x - 2 * y) * aLep * aLep * aLep;
  cs += ((7. / 12.) * b / x + (1. / 6.) * b / x / x - 0.5 * y) * aLep * aLep * aLep * aLep;
  cs *= 4 * kHc * kHc * 1e7 * kAlpha * kAlpha * M_PI / s;
  return cs;
}

double orc_dilep_cs_zm(int pdg, double aLep, double z, double m)
{
  const double mPart = lep_mass(pdg);
  const double s = m * m;
  const double k = sqrt(s) / 2.;
  const double p = sqrt(k * k - mPart * mPart);
  const double norm = 2 * M_PI * kAlpha * kAlpha / s * p / k;
  const double kt = -2 * k * (k - z * p) / mPart / mPart;
  const double ku = -2 * k * (k + z * p) / mPart / mPart;
  const double ks = kt + ku;
  const double kp = kt * ku;
  const double kq = 1. / kt + 1. / ku;
  const double kr = ku / kt + kt / ku;
  double cs = 0;
  cs += -8. * (4. * kq * kq + 4. * kq - kr);
  cs += 16. * (2. + kr) * aLep;
  cs += 4. * (2. - 4. * ks + kr) * aLep * aLep;
  cs += -8. * (2. + 2. * ks + kr) * aLep * aLep * aLep;
  cs += -4. * (4. + 2. * ks + 2. * kr - kp) * aLep * aLep * aLep * aLep;
  cs *= norm;
  return cs;
}

double orc_dilep_cs_m_pol_s(int pdg, double m)
{
  const double mPart = lep_mass(pdg);
  const double r = 2 * mPart / m;
  if (r > 1) return 0;
  return 4 * M_PI * kAlpha * kAlpha * kHc * kHc / m / m *
         ((1 + r * r - 3. / 4. * r * r * r * r) * 2 * log(1 / r + sqrt(1 / r / r - 1)) -
          (1 + 3. / 2. * r * r) * sqrt(1 - r * r));
}

double orc_dilep_cs_zm_pol_s(int pdg, double z, double m)
{
  const double mPart = lep_mass(pdg);
  const double mLep2 = mPart * mPart;
  const double m2 = m * m;
  const double z2 = z * z;
  double cs_s = 2 * M_PI * kAlpha * kAlpha;
  cs_s *= m2 - 4 * mLep2;
  cs_s *= sqrt(m2 - 4 * mLep2);
  cs_s *= (4 * mLep2 * (3 - 2 * z2 + z2 * z2)) + m2 * (1 - z2 * z2);
  cs_s /= m2 * m * (m2 * (1 - z2) + 4 * mLep2 * z2) * (m2 * (1 - z2) + 4 * mLep2 * z2);
  return cs_s;
}

double orc_dilep_cs_m_pol_ps(int pdg, double m)
{
  const double mPart = lep_mass(pdg);
  const double r = 2 * mPart / m;
  if (r > 1) return 0;
  return 4 * M_PI * kAlpha * kAlpha * kHc * kHc / m / m *
         ((1 + r * r - 1. / 4. * r * r * r * r) * 2 * log(1 / r + sqrt(1 / r / r - 1)) -
          (1 + 1. / 2. * r * r) * sqrt(1 - r * r));
}

double orc_dilep_cs_zm_pol_ps(int pdg, double z, double m)
{
  const double mPart = lep_mass(pdg);
  const double mLep2 = mPart * mPart;
  const double m2 = m * m;
  const double z2 = z * z;
  double cs_s = 2 * M_PI * kAlpha * kAlpha;
  cs_s *= sqrt(m2 - 4 * mLep2);
  cs_s *= m2 * m2 * (1 - z2 * z2) + 8 * m2 * mLep2 * (1 - z2 + z2 * z2) -
          16 * mLep2 * mLep2 * (1 - z2) * (1 - z2);
  cs_s /= m2 * m * (m2 * (1 - z2) + 4 * mLep2 * z2) * (m2 * (1 - z2) + 4 * mLep2 * z2);
  return cs_s;
}

/* src/UpcCrossSection.cpp:337-362 (parameter order of the definition) */
void orc_fill_cs_zm(const orc_ctx *c, double *out, int flag)
{
  const double scalingFactor = kHc * kHc * 1e7;
  const int nm = c->p.nm, nz = c->p.nz;
  const double dm = (c->p.mmax - c->p.mmin) / nm;
  const double dz = (c->p.zmax - c->p.zmin) / nz;
  const int pdg = c->p.proc_id;
  double cs = 0;
  for (int im = 0; im < nm; ++im) {
    const double m = c->p.mmin + dm * im;
    for (int iz = 0; iz < nz; ++iz) {
      const double z = c->p.zmin + dz * iz;
      if (flag == 0) cs = orc_dilep_cs_zm(pdg, c->p.lep_a, z, m);
      if (flag == 1) cs = orc_dilep_cs_zm_pol_s(pdg, z, m);
      if (flag == 2) cs = orc_dilep_cs_zm_pol_ps(pdg, z, m);
      out[(long)im * nz + iz] = cs * scalingFactor / dm;
    }
  }
}

/* src/UpcCrossSection.cpp:594-698, single-thread accumulation order (im-major) */
void orc_nuc_cs_ym(const orc_ctx *c, const double *lumi, const double *lumiS, const double *lumiP,
                   double *cs, double *ratio, double *tot_cs)
{
  const int nm = c->p.nm, ny = c->p.ny;
  const double dm = (c->p.mmax - c->p.mmin) / nm;
  const int pdg = c->p.proc_id;
  for (int im = 0; im < nm; ++im) {
    const double m = c->p.mmin + dm * im;
    for (int iy = 0; iy < ny; ++iy) {
      const long li = (long)im * ny + iy, oi = (long)iy * nm + im;
      if (!c->p.use_polarized) {
        cs[oi] = orc_dilep_cs_m(pdg, c->p.lep_a, m) * lumi[li];
      } else {
        const double cs_s = orc_dilep_cs_m_pol_s(pdg, m);
        const double cs_p = orc_dilep_cs_m_pol_ps(pdg, m);
        const double nuccs_s = lumiS[li] * cs_s;
        const double nuccs_p = lumiP[li] * cs_p;
        const double nuccs = nuccs_s + nuccs_p;
        cs[oi] = nuccs * 1e7;
        ratio[oi] = nuccs_s / nuccs_p;
      }
    }
  }
  for (int im = 0; im < nm; ++im)
    for (int iy = 0; iy < ny; ++iy) *tot_cs += cs[(long)iy * nm + im];
  *tot_cs *= 1e-6;
}

/* getPhotonPt's histogram (src/UpcCrossSection.cpp:1021-1051): bin contents
 * of the 5000-bin pt distribution over (0, 6 hc / R] at photon energy E. */
void orc_photon_pt_hist(const orc_ctx *c, double E, double gtot, double *out)
{
  const double pi2x4 = 4 * M_PI * M_PI;
  const double ereds = (E * E) / (gtot * gtot);
  const int nbins = 5000;
  for (int bin = 1; bin <= nbins; bin++) {
    const double pt = 6. * kHc / c->p.R / nbins * bin;
    const double arg = pt * pt + ereds;
    const double f = orc_spline_eval(&c->ff, arg, NULL);
    out[bin - 1] = (f * f) * pt * pt * pt / (pi2x4 * arg * arg);
  }
}

/* ------------------------------------------------------------------------ */
/* Histogram pdf (include/UpcSampler.h:53-71, 101-121)                       */
/* ------------------------------------------------------------------------ */
int orc_pdf_init(const double *bins, int n, double *sum)
{
  for (int i = 0; i < n; ++i)
    if (bins[i] < 0) return -1;
  double mean = 0, s = 0;
  for (int i = 0; i < n; ++i) mean += (bins[i] - mean) / ((double)(i + 1));
  sum[0] = 0;
  for (int i = 0; i < n; ++i) {
    s += (bins[i] / mean) / n;
    sum[i + 1] = s;
  }
  return 0;
}

double orc_pdf_sample(const double *sum, const double *range, int n, double r, int *bin)
{
  if (r == 1.0) r = 0.0;
  if (r < sum[0] || r >= sum[n]) { if (bin) *bin = -1; return NAN; }
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (hi + lo) / 2;
    if (r >= sum[mid]) lo = mid; else hi = mid;
  }
  if (bin) *bin = lo;
  const double delta = (r - sum[lo]) / (sum[lo + 1] - sum[lo]);
  return range[lo] + delta * (range[lo + 1] - range[lo]);
}

/* ------------------------------------------------------------------------ */
/* mt19937                                                                   */
/* ------------------------------------------------------------------------ */
void orc_mt_seed(orc_mt *s, unsigned long seed)
{
  if (seed == 0) seed = 4357; /* gsl_rng_set(r, 0) uses the generator's default seed */
  s->mt[0] = seed & 0xffffffffUL;
  for (int i = 1; i < 624; ++i) {
    s->mt[i] = (1812433253UL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (unsigned long)i);
    s->mt[i] &= 0xffffffffUL;
  }
  s->mti = 624;
}

static unsigned long mt_get(orc_mt *s)
{
  const unsigned long UPPER = 0x80000000UL, LOWER = 0x7fffffffUL;
  unsigned long k;
  if (s->mti >= 624) {
    int kk;
    for (kk = 0; kk < 624 - 397; ++kk) {
      const unsigned long y = (s->mt[kk] & UPPER) | (s->mt[kk + 1] & LOWER);
      s->mt[kk] = s->mt[kk + 397] ^ (y >> 1) ^ ((y & 1) ? 0x9908b0dfUL : 0UL);
    }
    for (; kk < 623; ++kk) {
      const unsigned long y = (s->mt[kk] & UPPER) | (s->mt[kk + 1] & LOWER);
      s->mt[kk] = s->mt[kk + (397 - 624)] ^ (y >> 1) ^ ((y & 1) ? 0x9908b0dfUL : 0UL);
    }
    {
      const unsigned long y = (s->mt[623] & UPPER) | (s->mt[0] & LOWER);
      s->mt[623] = s->mt[396] ^ (y >> 1) ^ ((y & 1) ? 0x9908b0dfUL : 0UL);
    }
    s->mti = 0;
  }
  k = s->mt[s->mti];
  k ^= (k >> 11);
  k ^= (k << 7) & 0x9d2c5680UL;
  k ^= (k << 15) & 0xefc60000UL;
  k ^= (k >> 18);
  s->mti++;
  return k & 0xffffffffUL;
}

double orc_mt_uniform(orc_mt *s) { return mt_get(s) / 4294967296.0; }
