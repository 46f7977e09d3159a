// Table builds of UpcCrossSection::init on the GPU (src/UpcCrossSection.cpp:116-163,
// 364-461, 752-1019): Woods-Saxon rho0, T_A and G_AA splines, the 10^6-knot
// form-factor spline and the Starlight breakup-probability spline.
#include "upc_internal.h"

#include <cstdio>

namespace upc {

// src/UpcCrossSection.cpp:139-150 evaluated on the fly: with even n the last
// weights are 2, 1 (kept as written).
template <class F>
__device__ double simpson_fn(int n, double h, F&& v)
{
  double sum = v(0) + v(n - 1);
  for (int i = 1; i < n - 1; i += 2) sum += 4. * v(i);
  for (int i = 2; i < n - 1; i += 2) sum += 2. * v(i);
  return sum * h / 3.;
}

// ---------------------------------------------------------------------------
// Natural cubic spline: LDL^T solve of the second-derivative system on the
// interior rows [w0, w1) (unknowns outside the window taken as 0), writing
// c[r+1] for rows r in [r0, r1).  With w0 = 0, w1 = m it is the exact global
// solve of gsl_interp_cspline; with a guard of G rows on each side of a chunk
// the truncation error is below (2-sqrt(3))^G of max|c|.
// ---------------------------------------------------------------------------
template <class XF, class YF>
__device__ void spline_window(XF&& xk, YF&& yk, int w0, int w1, int r0, int r1, double* c,
                              double* gam, double* zd, int stride)
{
  double off_prev = 0, gam_prev = 0, z_prev = 0;
#pragma unroll 4
  for (int r = w0; r < w1; ++r) {
    const double x0 = xk(r), x1 = xk(r + 1), x2 = xk(r + 2);
    const double y0 = yk(r), y1 = yk(r + 1), y2 = yk(r + 2);
    const double hi = x1 - x0, hip1 = x2 - x1;
    const double ydi = y1 - y0, ydip1 = y2 - y1;
    const double gi = hi != 0.0 ? 1.0 / hi : 0.0;
    const double gip1 = hip1 != 0.0 ? 1.0 / hip1 : 0.0;
    const double diag = 2.0 * (hip1 + hi);
    const double g = 3.0 * (ydip1 * gip1 - ydi * gi);
    double alp, z;
    if (r == w0) {
      alp = diag;
      z = g;
    } else {
      alp = diag - off_prev * gam_prev;
      z = g - gam_prev * z_prev;
    }
    const double gm = hip1 / alp;
    gam[(size_t)(r - w0) * stride] = gm;
    zd[(size_t)(r - w0) * stride] = z / alp;
    off_prev = hip1;
    gam_prev = gm;
    z_prev = z;
  }
  double cnext = zd[(size_t)(w1 - 1 - w0) * stride];
  if (w1 - 1 >= r0 && w1 - 1 < r1) c[w1] = cnext;
#pragma unroll 4
  for (int r = w1 - 2; r >= w0; --r) {
    const double cr = zd[(size_t)(r - w0) * stride] - gam[(size_t)(r - w0) * stride] * cnext;
    if (r >= r0 && r < r1) c[r + 1] = cr;
    cnext = cr;
  }
}

__device__ __forceinline__ double4 spline_coef(double xl, double xh, double yl, double yh, double ci,
                                               double cip1)
{
  const double h = xh - xl;
  const double dy = yh - yl;
  const double bi = dy / h - h * (cip1 + 2.0 * ci) / 3.0;
  const double di = (cip1 - ci) / (3.0 * h);
  return make_double4(yl, bi, ci, di);
}

// ---------------------------------------------------------------------------
// calcWSRho, src/UpcCrossSection.cpp:152-163
// ---------------------------------------------------------------------------
__global__ void k_ws_rho0(double R, double a, int A, DevScalars* sc)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double bmax = 20.;
  const double db = bmax / (kNB - 1.);
  const double s = simpson_fn(kNB, db, [&](int ib) {
    const double r = ib * db;
    return r * r / (1. + exp((r - R) / a));
  });
  sc->rho0 = A / s / 4. / kPi;
}

// ---------------------------------------------------------------------------
// prepareGAA, src/UpcCrossSection.cpp:364-414, in three launches: T_A and its
// spline (one block); the G_AA integral of each b (one block per b: the 200
// integrand values in parallel, then the quirky Simpson sum in the reference's
// order by one thread); the G_AA spline (one block).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double gaa_cs_nn(double sqrts)
{
  const double ssm = pow(sqrts, 2) / pow(2 * kMProt + 2.1206, 2);
  return 0.1 * (34.41 + 0.2720 * pow(log(ssm), 2) + 13.07 * pow(ssm, -0.4473) - 7.394 * pow(ssm, -0.5486));
}

__global__ __launch_bounds__(256) void k_gaa_ta(double R, double a, double sqrts, DevScalars* sc,
                                                double4* out_ta)
{
  __shared__ double vTA[kNB], cc[kNB], gam[kNB], zd[kNB];
  const double db = 20. / (kNB - 1);
  const int t = threadIdx.x;
  const double rho0 = sc->rho0;
  if (t < kNB) {
    const double b = t * db;
    vTA[t] = 2. * simpson_fn(kNB, db, [&](int iz) {
      const double z = iz * db;
      const double r = sqrt(b * b + z * z);
      return rho0 / (1 + exp((r - R) / a));
    });
  }
  __syncthreads();
  auto xk = [&](int i) { return i * db; };
  if (t == 0) {
    cc[0] = 0.0;
    cc[kNB - 1] = 0.0;
    spline_window(xk, [&](int i) { return vTA[i]; }, 0, kNB - 2, 0, kNB - 2, cc, gam, zd, 1);
    sc->csNN = gaa_cs_nn(sqrts);
  }
  __syncthreads();
  if (t < kNB - 1) out_ta[t] = spline_coef(xk(t), xk(t + 1), vTA[t], vTA[t + 1], cc[t], cc[t + 1]);
}

__global__ __launch_bounds__(256) void k_gaa_g(const DevScalars* sc, const double4* __restrict__ ta_coef,
                                               double* __restrict__ vG)
{
  __shared__ double4 coefTA[kNB];
  __shared__ double v[kNB], cosk[5];
  const double bmax = 20.;
  const double db = bmax / (kNB - 1);
  const int t = threadIdx.x;
  if (t < 5) cosk[t] = cos(kPi * kGLx5[t]);
  if (t < kNB - 1) coefTA[t] = ta_coef[t];
  __syncthreads();
  const double inv_db = 1.0 / db;
  auto ta = [&](double r) { return spline_eval(coefTA, 0.0, db, inv_db, kNB, r); };
  const double b = blockIdx.x * db;
  if (t < kNB) {
    const double s = t * db;
    double sum_phi = 0;
    for (int k = 0; k < 5; ++k) {
      const double r = sqrt(b * b + s * s + 2. * b * s * cosk[k]);
      sum_phi += 2. * kPi * kGLw5[k] * ta(r < bmax ? r : bmax);
    }
    v[t] = 2. * s * ta(s < bmax ? s : bmax) * sum_phi;
  }
  __syncthreads();
  if (t == 0) vG[blockIdx.x] = exp(-sc->csNN * simpson_fn(kNB, db, [&](int is) { return v[is]; }));
}

__global__ __launch_bounds__(256) void k_gaa_spline(const double* __restrict__ vG_g, double4* out_gaa)
{
  __shared__ double vG[kNB], cc[kNB], gam[kNB], zd[kNB];
  const double db = 20. / (kNB - 1);
  const int t = threadIdx.x;
  if (t < kNB) vG[t] = vG_g[t];
  __syncthreads();
  auto xk = [&](int i) { return i * db; };
  if (t == 0) {
    cc[0] = 0.0;
    cc[kNB - 1] = 0.0;
    spline_window(xk, [&](int i) { return vG[i]; }, 0, kNB - 2, 0, kNB - 2, cc, gam, zd, 1);
  }
  __syncthreads();
  if (t < kNB - 1) out_gaa[t] = spline_coef(xk(t), xk(t + 1), vG[t], vG[t + 1], cc[t], cc[t + 1]);
}

// ---------------------------------------------------------------------------
// calcFormFac / prepareFormFac, src/UpcCrossSection.cpp:436-461
// ---------------------------------------------------------------------------
__device__ __forceinline__ double calc_form_fac(double Q2, double R, double a, double rho0)
{
  const double Q = sqrt(Q2) / kHc;
  const double coshVal = cosh(kPi * Q * a);
  const double sinhVal = sinh(kPi * Q * a);
  double ff = 4 * kPi * kPi * rho0 * a * a * a / (Q * a * Q * a * sinhVal * sinhVal) *
              (kPi * Q * a * coshVal * sin(Q * R) - Q * R * cos(Q * R) * sinhVal);
  ff += 8 * kPi * rho0 * a * a * a * exp(-R / a) / (1 + Q * Q * a * a) / (1 + Q * Q * a * a);
  return ff;
}

__global__ void k_ff_values(double R, double a, const DevScalars* sc, double* y)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNQ2) return;
  y[i] = calc_form_fac(knot(kQ2min, kDQ2, i), R, a, sc->rho0);
}

// Windowed natural-spline solve on uniform knots x_i = x0 + i*dx.
__global__ void k_spline_window(double x0, double dx, const double* __restrict__ y, int n,
                                double* __restrict__ c, double* __restrict__ scratch, int chunk,
                                int guard, int nthr)
{
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = n - 2;
  if (tid == 0) {
    c[0] = 0.0;
    c[n - 1] = 0.0;
  }
  const int r0 = tid * chunk;
  if (tid >= nthr || r0 >= m) return;
  const int r1 = min(m, r0 + chunk);
  const int w0 = max(0, r0 - guard);
  const int w1 = min(m, r1 + guard);
  const int wlen = chunk + 2 * guard;
  double* gam = scratch + tid;
  double* zd = scratch + (size_t)wlen * nthr + tid;
  spline_window([&](int i) { return knot(x0, dx, i); }, [&](int i) { return y[i]; }, w0, w1, r0, r1,
                c, gam, zd, nthr);
}

__global__ void k_spline_coefs(double x0, double dx, const double* __restrict__ y,
                               const double* __restrict__ c, int n, double4* __restrict__ coef)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  coef[i] = spline_coef(knot(x0, dx, i), knot(x0, dx, i + 1), y[i], y[i + 1], c[i], c[i + 1]);
}

// ---------------------------------------------------------------------------
// calcBreakupProb, src/UpcCrossSection.cpp:752-1019 (Starlight).  Its data
// tables (Pb GDR, Lepretre, Carlos, Armstrong, Michalowski) are reproduced
// as written; Pb is hardcoded for every nucleus there and here.
// ---------------------------------------------------------------------------
__device__ const double kBkE1[23] = {0., 103., 106., 112., 119., 127., 132., 145., 171., 199., 230., 235.,
                                     254., 280., 300., 320., 330., 333., 373., 390., 420., 426., 440.};
__device__ const double kBkS1[23] = {0., 12.0, 11.5, 12.0, 12.0, 12.0, 15.0, 17.0, 28.0, 33.0,
                                     52.0, 60.0, 70.0, 76.0, 85.0, 86.0, 89.0, 89.0, 75.0, 76.0, 69.0, 59.0, 61.0};
__device__ const double kBkE2[12] = {0., 2000., 3270., 4100., 4810., 6210., 6600.,
                                     7790., 8400., 9510., 13600., 16400.};
__device__ const double kBkS2[12] = {0., .1266, .1080, .0805, .1017, .0942, .0844, .0841, .0755, .0827,
                                     .0626, .0740};
__device__ const double kBkE3[29] = {0., 26., 28., 30., 32., 34., 36., 38., 40., 44., 46., 48., 50., 52., 55.,
                                     57., 62., 64., 66., 69., 72., 74., 76., 79., 82., 86., 92., 98., 103.};
__device__ const double kBkS3[29] = {0., 30., 21.5, 22.5, 18.5, 17.5, 15., 14.5, 19., 17.5, 16., 14.,
                                     20., 16.5, 17.5, 17., 15.5, 18., 15.5, 15.5, 15., 13.5, 18., 14.5, 15.5, 12.5, 13.,
                                     13., 12.};
__device__ const double kBkSigt[160] = {0., .4245, .4870, .5269, .4778, .4066, .3341, .2444, .2245, .2005,
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
__device__ const double kBkSigtn[160] = {0., .3125, .3930, .4401, .4582, .3774, .3329, .2996, .2715, .2165,
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

// First-call initialisation of the calcBreakupProb statics for Pb
// (src/UpcCrossSection.cpp:866-969); the one-neutron tables (eee, sa) feed
// only the unused p1n sum and are not built.
// The photo-absorption tables of calcBreakupProb's first call
// (src/UpcCrossSection.cpp:866-969) for this context's gamma.  Entries are
// independent except the Regge energies s_j = s_{j-1} * 10^0.05, which one
// thread chains in the reference's order; the rest is filled in parallel.
__global__ __launch_bounds__(256) void k_breakup_init(double g1, double* ee, double* se, DevScalars* sc)
{
  __shared__ double sreg[160];
  const int t = threadIdx.x;
  const double gammatarg = 2. * g1 * g1 - 1.;
  const double hbarcmev = 197.3269718;
  const double pi = 3.14159;
  const int zp = 82, ap = 208;
  const double si1 = 640., gw = 4.05, o1 = 13.42, o0 = 7.4;
  for (int i = t; i < kMaxEE; i += blockDim.x) {
    ee[i] = 0.;
    se[i] = 0.;
  }
  const double delo = .05;
  const double scon = .1 * gw * gw * si1;
  const double zcon = zp / (gammatarg * (pi) * (hbarcmev)) * zp / (gammatarg * (pi) * (hbarcmev)) / 137.04;
  const int ne = int((25. - o0) / delo) + 1;
  const double x = .0677, y = .129, eps = .0808, eta = .4525, em = .94;
  int ictr = 100;
  if (gammatarg > (2. * 150. * 150.)) ictr = 150;
  // index layout: GDR 1..ne, then 27 + 22 + 62 + 11 tabulated, then ictr Regge, then the sentinel
  const int i3 = ne, i1 = i3 + 27, i25 = i1 + 22, i2 = i25 + 62, ir = i2 + 11;
  __syncthreads();  // zeroing done before the entries are written
  for (int i = 1 + t; i <= ne; i += blockDim.x) {
    const double e = o0 + (i - 1) * delo;
    ee[i] = e;
    se[i] = scon * e * e / (((o1 * o1 - e * e) * (o1 * o1 - e * e)) + e * e * gw * gw);
  }
  for (int j = 1 + t; j <= 27; j += blockDim.x) {
    ee[i3 + j] = kBkE3[j];
    se[i3 + j] = .1 * ap * kBkS3[j] / 208.;
  }
  for (int j = 1 + t; j <= 22; j += blockDim.x) {
    ee[i1 + j] = kBkE1[j];
    se[i1 + j] = .1 * ap * kBkS1[j] / 208.;
  }
  // ee = previous + 25 from kBkE1[22] = 440: integers, exact in any order
  for (int j = 9 + t; j <= 70; j += blockDim.x) {
    ee[i25 + j - 8] = kBkE1[22] + 25. * (j - 8);
    se[i25 + j - 8] = .1 * (zp * kBkSigt[j] + (ap - zp) * kBkSigtn[j]);
  }
  for (int j = 1 + t; j <= 11; j += blockDim.x) {
    ee[i2 + j] = kBkE2[j];
    se[i2 + j] = .1 * ap * kBkS2[j];
  }
  if (t == 0) {
    const double exx = pow(10., .05);
    double sv = .002 * em * kBkE2[11];  // ee[ir] = the last E2 energy
    for (int j = 1; j <= ictr; ++j) {
      sv = sv * exx;
      sreg[j - 1] = sv;
    }
  }
  __syncthreads();
  for (int j = 1 + t; j <= ictr; j += blockDim.x) {
    const double sv = sreg[j - 1];
    ee[ir + j] = 1000. * .5 * (sv - em * em) / em;
    const double pom = x * pow(sv, eps);
    const double vec = y * pow(sv, (-eta));
    se[ir + j] = .1 * .65 * ap * (pom + vec);
  }
  if (t == 0) {
    ee[ir + ictr + 1] = 99999999999.;
    sc->zcon = zcon;
    sc->o0 = o0;
    sc->nee = ir + ictr + 1;
  }
}

// Per-b Coulomb breakup probability (src/UpcCrossSection.cpp:972-1018).
__device__ __forceinline__ double breakup_prob(double b, int mode, double g1, const double* ee,
                                               const double* se, double zcon, double o0)
{
  const double hbarcmev = 197.3269718;
  const double gammatarg = 2. * g1 * g1 - 1.;
  const double omaxx = g1 > 500. ? 1.E10 : 1.E7;
  double pxn = 0.;
  const double omax = fmin(omaxx, 4. * gammatarg * (hbarcmev) / b);
  if (omax < o0) return 0.;
  const double scale = 1.0 / ((hbarcmev)*gammatarg);
  double gk1m = bessel_k1_as(ee[1] * b * scale);
  int k = 2;
  while (ee[k] < omax) {
    const double gk1 = bessel_k1_as(ee[k] * b * scale);
    pxn = pxn + zcon * (ee[k] - ee[k - 1]) * .5 * (se[k - 1] * ee[k - 1] * gk1m * gk1m + se[k] * ee[k] * gk1 * gk1);
    k = k + 1;
    gk1m = gk1;
  }
  double P = 0.;
  if (mode == 1) P = 1.;
  if (mode == 2) P = (1 - exp(-1 * pxn)) * (1 - exp(-1 * pxn));
  if (mode == 3) P = exp(-2 * pxn);
  if (mode == 4) P = 2. * exp(-pxn) * (1. - exp(-pxn));
  return P;
}

// P(b) at the breakup knots (d_b == nullptr) or at caller points.  The
// trapezoid sum over the ~600 photon energies of one b is split over a group
// of kBkLanes consecutive lanes (contiguous energy ranges, partial sums added
// by shuffles): the same terms, summed in another order.
constexpr int kBkLanes = 8;

__global__ __launch_bounds__(256) void k_breakup_values(double g1, int mode, const double* __restrict__ ee_g,
                                                        const double* __restrict__ se_g,
                                                        const DevScalars* sc, int n,
                                                        const double* __restrict__ d_b,
                                                        double* __restrict__ out)
{
  __shared__ double ee[kMaxEE], se[kMaxEE];
  const int nee = sc->nee;
  for (int i = threadIdx.x; i <= nee && i < kMaxEE; i += blockDim.x) {
    ee[i] = ee_g[i];
    se[i] = se_g[i];
  }
  __syncthreads();
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int sub = gid % kBkLanes;
  const int i0 = gid / kBkLanes;
  const bool live = i0 < n;
  const int i = live ? i0 : n - 1;
  const double b = d_b ? d_b[i] : knot(kBkMin, kDBk, i);
  // src/UpcCrossSection.cpp:972-1018, as breakup_prob()
  const double hbarcmev = 197.3269718;
  const double gammatarg = 2. * g1 * g1 - 1.;
  const double omaxx = g1 > 500. ? 1.E10 : 1.E7;
  const double omax = fmin(omaxx, 4. * gammatarg * (hbarcmev) / b);
  const double scale = 1.0 / ((hbarcmev)*gammatarg);
  const double zcon = sc->zcon;
  // terms k = 2 .. kend - 1, kend = the first k with ee[k] >= omax (ee ascends)
  int kend = 2;
  {
    int lo = 2, hi = nee;  // ee[nee] is the 1e11 sentinel
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ee[mid] < omax) lo = mid + 1;
      else hi = mid;
    }
    kend = lo;
  }
  const int nterm = kend - 2;
  const int per = (nterm + kBkLanes - 1) / kBkLanes;
  const int ka = 2 + sub * per, kb = min(kend, ka + per);
  double part = 0.;
  if (ka < kb) {
    double gk1m = bessel_k1_as(ee[ka - 1] * b * scale);
    for (int k = ka; k < kb; ++k) {
      const double gk1 = bessel_k1_as(ee[k] * b * scale);
      part = part + zcon * (ee[k] - ee[k - 1]) * .5 * (se[k - 1] * ee[k - 1] * gk1m * gk1m + se[k] * ee[k] * gk1 * gk1);
      gk1m = gk1;
    }
  }
#pragma unroll
  for (int o = 1; o < kBkLanes; o <<= 1) part += __shfl_xor(part, o);
  if (!live || sub != 0) return;
  const double pxn = part;
  double P = 0.;
  if (omax >= sc->o0) {
    if (mode == 1) P = 1.;
    if (mode == 2) P = (1 - exp(-1 * pxn)) * (1 - exp(-1 * pxn));
    if (mode == 3) P = exp(-2 * pxn);
    if (mode == 4) P = 2. * exp(-pxn) * (1. - exp(-pxn));
  }
  out[i] = P;
}

// p20 and bskip, one wave.  bskip is the knot before the first G_AA interval
// whose bound |y| + |b| h + |c| h^2 + |d| h^3 (h a little over the knot
// spacing, so rounding of the interval estimate is covered) reaches
// p20 2^-71: below it every evaluation of the G_AA spline, on whichever
// interval the estimate names, is smaller than p20 2^-70 in magnitude.
__global__ __launch_bounds__(64) void k_p20(const double4* __restrict__ bk, const double4* __restrict__ gaa, int mode,
                                            DevScalars* sc)
{
  const int lane = threadIdx.x;
  const double p20 = mode == 1 ? 1.0 : spline_eval(bk, kBkMin, kDBk, 1.0 / kDBk, kNBkUsed, 20.);
  const double db = 20. / (kNB - 1), h = db * 1.001;
  const double lim = p20 * 0x1p-71;
  int first = kNB - 1;
  for (int i = lane; i < kNB - 1; i += 64) {
    const double4 c = gaa[i];
    const double bound = fabs(c.x) + h * (fabs(c.y) + h * (fabs(c.z) + h * fabs(c.w)));
    if (!(bound < lim)) first = min(first, i);
  }
  for (int o = 32; o > 0; o >>= 1) first = min(first, __shfl_xor(first, o));
  if (lane == 0) {
    sc->p20 = p20;
    sc->bskip = first > 0 ? (first - 1) * db : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Evaluation kernels for the batched C-ABI functions
// ---------------------------------------------------------------------------
__global__ void k_eval_ff(const double4* __restrict__ ff, double R, double a, const DevScalars* sc,
                          const double* __restrict__ q2, int n, int tabulated, double* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = q2[i];
  if (tabulated) {
    if (!(v >= kQ2min && v <= knot(kQ2min, kDQ2, kNQ2 - 1))) {
      out[i] = __builtin_nan("");  // outside the table: GSL_EDOM in the reference
      return;
    }
    out[i] = spline_eval(ff, kQ2min, kDQ2, 1.0 / kDQ2, kNQ2, v);
  } else {
    out[i] = calc_form_fac(v, R, a, sc->rho0);
  }
}

__global__ void k_eval_bk_spline(const double4* __restrict__ bk, const double* __restrict__ b, int n,
                                 double* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = b[i];
  const double hi = knot(kBkMin, kDBk, kNBkUsed - kBkGuard);
  if (!(v >= kBkMin && v <= hi)) {
    out[i] = __builtin_nan("");
    return;
  }
  out[i] = spline_eval(bk, kBkMin, kDBk, 1.0 / kDBk, kNBkUsed, v);
}

__global__ void k_eval_overlap(const double4* __restrict__ tab, const double* __restrict__ b, int n,
                               double* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double db = 20. / (kNB - 1);
  const double v = b[i];
  if (!(v >= 0.0 && v <= 20.0)) {
    out[i] = __builtin_nan("");
    return;
  }
  out[i] = spline_eval(tab, 0.0, db, 1.0 / db, kNB, v);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Rows solved per thread, and the guard rows of its window on each side
// ((2 - sqrt 3)^48 = 3e-28): short serial chains, 15.6 k threads for 10^6 knots.
constexpr int kSplineChunk = 64;
constexpr int kSplineGuard = 48;

int alloc_tables(upc_ctx* c)
{
  UPC_HIP(hipMalloc(&c->d_scalars, sizeof(DevScalars)));
  UPC_HIP(hipMalloc(&c->d_ta, sizeof(double4) * kNB));
  UPC_HIP(hipMalloc(&c->d_gaa, sizeof(double4) * kNB));
  UPC_HIP(hipMalloc(&c->d_ff_y, sizeof(double) * kNQ2));
  UPC_HIP(hipMalloc(&c->d_ff_c, sizeof(double) * kNQ2));
  UPC_HIP(hipMalloc(&c->d_ff, sizeof(double4) * kNQ2));
  UPC_HIP(hipMalloc(&c->d_bk_y, sizeof(double) * kNBkUsed));
  UPC_HIP(hipMalloc(&c->d_bk_c, sizeof(double) * kNBkUsed));
  UPC_HIP(hipMalloc(&c->d_bk, sizeof(double4) * kNBkUsed));
  UPC_HIP(hipMalloc(&c->d_ee, sizeof(double) * kMaxEE));
  UPC_HIP(hipMalloc(&c->d_se, sizeof(double) * kMaxEE));
  c->spline_scratch_elems = (size_t)2 * (kSplineChunk + 2 * kSplineGuard) * nblk(kNQ2, kSplineChunk);
  UPC_HIP(hipMalloc(&c->d_spline_scratch, sizeof(double) * c->spline_scratch_elems));
  c->bk_scratch_elems = (size_t)2 * (kSplineChunk + 2 * kSplineGuard) * nblk(kNBkUsed, kSplineChunk);
  UPC_HIP(hipMalloc(&c->d_bk_scratch, sizeof(double) * c->bk_scratch_elems));
  UPC_HIP(hipMalloc(&c->d_gaa_scratch, sizeof(double) * kNB));
  for (int i = 0; i < 2; ++i) {
    UPC_HIP(hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking));
    UPC_HIP(hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming));
  }
  UPC_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  return 0;
}

void free_tables(upc_ctx* c)
{
  for (int i = 0; i < 2; ++i) {
    if (c->side[i]) {
      (void)hipStreamSynchronize(c->side[i]);
      (void)hipStreamDestroy(c->side[i]);
    }
    if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  void* ptrs[] = {c->d_scalars, c->d_ta, c->d_gaa, c->d_ff_y, c->d_ff_c, c->d_ff,
                  c->d_bk_y, c->d_bk_c, c->d_bk, c->d_ee, c->d_se, c->d_spline_scratch,
                  c->d_bk_scratch, c->d_gaa_scratch};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

static int solve_spline(double x0, double dx, const double* y, int n, double* cc, double4* coef,
                        double* scratch, size_t scratch_elems, hipStream_t s)
{
  const int m = n - 2;
  const int nthr = nblk(m, kSplineChunk);
  if ((size_t)2 * (kSplineChunk + 2 * kSplineGuard) * nthr > scratch_elems) {
    set_error("spline scratch too small");
    return UPC_EINVAL;
  }
  hipLaunchKernelGGL(k_spline_window, dim3(nblk(nthr, 64)), dim3(64), 0, s, x0, dx, y, n, cc, scratch,
                     kSplineChunk, kSplineGuard, nthr);
  hipLaunchKernelGGL(k_spline_coefs, dim3(nblk(n - 1, 256)), dim3(256), 0, s, x0, dx, y, cc, n, coef);
  UPC_HIP(hipGetLastError());
  return 0;
}

int build_tables(upc_ctx* c, hipStream_t s)
{
  const upc_config& p = c->cfg;
  {
  PhaseScope ph(c, s, UPC_PHASE_TABLES);
  // rho0 feeds T_A and the form factor; then three independent chains, each
  // with its own scratch and its own fields of DevScalars: G_AA (side[0]),
  // breakup (side[1]) and the form factor (s), joined before k_p20, which
  // reads both splines.
  hipLaunchKernelGGL(k_ws_rho0, dim3(1), dim3(64), 0, s, p.R, p.a, p.A, c->d_scalars);
  UPC_HIP(hipEventRecord(c->ev_fork, s));
  const hipStream_t sg = c->side[0], sb = c->side[1];
  UPC_HIP(hipStreamWaitEvent(sg, c->ev_fork, 0));
  UPC_HIP(hipStreamWaitEvent(sb, c->ev_fork, 0));
  hipLaunchKernelGGL(k_gaa_ta, dim3(1), dim3(256), 0, sg, p.R, p.a, p.sqrts, c->d_scalars, c->d_ta);
  hipLaunchKernelGGL(k_gaa_g, dim3(kNB), dim3(256), 0, sg, c->d_scalars, c->d_ta, c->d_gaa_scratch);
  hipLaunchKernelGGL(k_gaa_spline, dim3(1), dim3(256), 0, sg, c->d_gaa_scratch, c->d_gaa);
  UPC_HIP(hipGetLastError());
  UPC_HIP(hipEventRecord(c->ev_join[0], sg));
  hipLaunchKernelGGL(k_breakup_init, dim3(1), dim3(256), 0, sb, c->g1, c->d_ee, c->d_se, c->d_scalars);
  int rc = 0;
  if (p.breakup_mode > 1) {
    hipLaunchKernelGGL(k_breakup_values, dim3(nblk((long)kNBkUsed * kBkLanes, 256)), dim3(256), 0, sb, c->g1,
                       p.breakup_mode, c->d_ee, c->d_se, c->d_scalars, kNBkUsed, (const double*)nullptr,
                       c->d_bk_y);
    UPC_HIP(hipGetLastError());
    rc = solve_spline(kBkMin, kDBk, c->d_bk_y, kNBkUsed, c->d_bk_c, c->d_bk, c->d_bk_scratch,
                      c->bk_scratch_elems, sb);
  }
  UPC_HIP(hipEventRecord(c->ev_join[1], sb));
  hipLaunchKernelGGL(k_ff_values, dim3(nblk(kNQ2, 256)), dim3(256), 0, s, p.R, p.a, c->d_scalars, c->d_ff_y);
  UPC_HIP(hipGetLastError());
  if (!rc)
    rc = solve_spline(kQ2min, kDQ2, c->d_ff_y, kNQ2, c->d_ff_c, c->d_ff, c->d_spline_scratch,
                      c->spline_scratch_elems, s);
  UPC_HIP(hipStreamWaitEvent(s, c->ev_join[0], 0));
  UPC_HIP(hipStreamWaitEvent(s, c->ev_join[1], 0));
  if (rc) return rc;
  hipLaunchKernelGGL(k_p20, dim3(1), dim3(64), 0, s, c->d_bk, c->d_gaa, p.breakup_mode, c->d_scalars);
  UPC_HIP(hipGetLastError());
  }
  UPC_HIP(hipMemcpyAsync(c->hp_scalars, c->d_scalars, sizeof(DevScalars), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipEventRecord(c->ev_scalars, s));
  c->scalars_pending = true;
  return 0;
}

int launch_form_factor(upc_ctx* c, const double* d_q2, int n, int tabulated, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_eval_ff, dim3(nblk(n, 256)), dim3(256), 0, s, c->d_ff, c->cfg.R, c->cfg.a,
                     c->d_scalars, d_q2, n, tabulated, d_out);
  UPC_HIP(hipGetLastError());
  return 0;
}

int launch_breakup(upc_ctx* c, const double* d_b, int n, int mode, int tabulated, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  if (tabulated) {
    if (c->cfg.breakup_mode <= 1) {
      set_error("breakup table exists only for BREAKUP_MODE > 1");
      return UPC_ESTATE;
    }
    hipLaunchKernelGGL(k_eval_bk_spline, dim3(nblk(n, 256)), dim3(256), 0, s, c->d_bk, d_b, n, d_out);
  } else {
    if (mode < 1 || mode > 4) {
      set_error("breakup mode must be 1..4");
      return UPC_EINVAL;
    }
    hipLaunchKernelGGL(k_breakup_values, dim3(nblk((long)n * kBkLanes, 256)), dim3(256), 0, s, c->g1, mode,
                       c->d_ee, c->d_se, c->d_scalars, n, d_b, d_out);
  }
  UPC_HIP(hipGetLastError());
  return 0;
}

int launch_overlap(upc_ctx* c, const double* d_b, int n, int which, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_eval_overlap, dim3(nblk(n, 256)), dim3(256), 0, s, which ? c->d_ta : c->d_gaa,
                     d_b, n, d_out);
  UPC_HIP(hipGetLastError());
  return 0;
}

}  // namespace upc
