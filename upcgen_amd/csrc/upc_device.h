// Device numerics for the upcgen luminosity path on CDNA4 (gfx950).
//
// Everything here is FP64 and runs on the GPU.  Functions restate the
// arithmetic of the reference (upcgen src/UpcCrossSection.cpp) and of the
// third-party routines it calls:
//   gsl_sf_bessel_J1/K0/K1 -> Chebyshev / power series generated from the
//                             mathematics (gen_coeffs.py, DLMF 10.17, 10.31)
//   TMath::BesselK1/I1     -> Abramowitz & Stegun 9.8.3, 9.8.4, 9.8.7, 9.8.8
//   gsl_interp_cspline     -> natural cubic spline on uniform knots
//   gsl_integration_qags   -> QUADPACK dqagse (QK21 + Wynn epsilon)
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#include "upc_coeffs.h"

namespace upc {

// include/UpcPhysConstants.h:26-32
constexpr double kAlpha = 1.0 / 137.035999074;
constexpr double kHc = 0.1973269718;
constexpr double kMProt = 0.9382720813;
constexpr double kMNeut = 0.939565346;
constexpr double kMEl = 0.000510998946;
constexpr double kMMu = 0.1056583745;
constexpr double kMTau = 1.77686;
constexpr double kPi = 3.14159265358979323846;

// include/UpcCrossSection.h:90-111, the 5 non-negative abscissas of the
// 10-point Gauss-Legendre table with their weights, in the reference's order.
__device__ constexpr double kGLx5[5] = {0.1488743389816312, 0.4333953941292472, 0.6794095682990244,
                                        0.8650633666889845, 0.9739065285171717};
__device__ constexpr double kGLw5[5] = {0.2955242247147529, 0.2692667193099963, 0.2190863625159820,
                                        0.1494513491505806, 0.0666713443086881};

// src/UpcCrossSection.cpp:46-49: form-factor table grid
constexpr double kQ2min = 1e-9;
constexpr double kQ2max = 2.;
constexpr int kNQ2 = 1000000;
constexpr double kDQ2 = (kQ2max - kQ2min) / kNQ2;
// Last knot of the table, Q2min + (nQ2 - 1) dQ2 with each operation rounded
// (as the reference's table loop computes it): the end of the spline's range.
constexpr double kQ2Last = kQ2min + (double)(kNQ2 - 1) * kDQ2;

// src/UpcCrossSection.cpp:418-421: breakup table grid
constexpr double kBkMin = 1e-6;
constexpr double kBkMax = 1000;
constexpr int kNBk = 1000000;
constexpr double kDBk = (kBkMax - kBkMin) / kNBk;

constexpr int kNB1 = 120, kNB2 = 120, kNB = 200;

// ---------------------------------------------------------------------------
// Chebyshev / polynomial evaluation
// ---------------------------------------------------------------------------
// Coefficients live in __constant__ memory and are read with uniform (scalar)
// loads inside partly unrolled loops (4 terms per trip: one scalar-load wait
// per trip); fully unrolled loops of FP64 immediates get hoisted into ~2
// VGPRs per coefficient by the loop optimiser.
template <int N>
__device__ __forceinline__ double clenshaw(const double (&c)[N], double u)
{
  double b1 = 0.0, b2 = 0.0;
  const double u2 = 2.0 * u;
#pragma unroll
  for (int k = N - 1; k >= 1; --k) {
    const double t = fma(u2, b1, c[k] - b2);
    b2 = b1;
    b1 = t;
  }
  return fma(u, b1, c[0] - b2);
}

// Two Chebyshev series of the same length in the same variable.
template <int N>
__device__ __forceinline__ void clenshaw2(const double (&c)[N], const double (&d)[N], double u, double* fc,
                                          double* fd)
{
  double b1 = 0.0, b2 = 0.0, e1 = 0.0, e2 = 0.0;
  const double u2 = 2.0 * u;
#pragma unroll 4
  for (int k = N - 1; k >= 1; --k) {
    const double t = fma(u2, b1, c[k] - b2);
    b2 = b1;
    b1 = t;
    const double s = fma(u2, e1, d[k] - e2);
    e2 = e1;
    e1 = s;
  }
  *fc = fma(u, b1, c[0] - b2);
  *fd = fma(u, e1, d[0] - e2);
}

template <int N>
__device__ __forceinline__ double horner(const double (&c)[N], double y)
{
  double s = c[N - 1];
#pragma unroll 1
  for (int k = N - 2; k >= 0; --k) s = c[k] + y * s;
  return s;
}

// sin and cos of |x| < 1e5: Cody-Waite reduction by pi/2 split in three
// doubles (each step one FMA), then near-minimax polynomials in r^2 on
// |r| <= pi/4 (kSinPoly, kCosPoly: absolute errors 1.4e-17 and 9e-19).
// a + b c as one three-address v_fma_f64.  For a Horner step with a
// loop-invariant coefficient `a` the compiler otherwise emits v_fmac_f64
// (accumulator tied to `a`) preceded by a copy of `a`.
__device__ __forceinline__ double fma3(double b, double c, double a)
{
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(c), "s"(a));
  return r;
}

__device__ __forceinline__ void sincos_medium(double x, double* s, double* c)
{
  constexpr double kTwoOverPi = 0.6366197723675814;
  constexpr double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.497384905065873e-33;
  const double n = rint(x * kTwoOverPi);
  double r = fma(-n, P1, x);
  r = fma(-n, P2, r);
  r = fma(-n, P3, r);
  const double r2 = r * r;
  double ps = kSinPoly[kSinPoly_N - 1];
#pragma unroll
  for (int k = kSinPoly_N - 2; k >= 0; --k) ps = fma3(r2, ps, kSinPoly[k]);
  const double sr = fma(r * r2, ps, r);
  double pc = kCosPoly[kCosPoly_N - 1];
#pragma unroll
  for (int k = kCosPoly_N - 2; k >= 0; --k) pc = fma3(r2, pc, kCosPoly[k]);
  const double cr = fma(r2 * r2, pc, fma(-0.5, r2, 1.0));
  const int q = (int)n & 3;  // |n| < 2^31 for |x| < 1e5
  const double ss = (q & 1) ? cr : sr;
  const double cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// 1/x for normal x: the hardware estimate refined by two Newton steps
// r += r (1 - x r) (each squares the relative error).
__device__ __forceinline__ double rcp_nr(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// 1/sqrt(x) for normal x > 0: the hardware estimate refined by two Newton
// steps y += y (1/2 - (x/2) y^2) (each squares the relative error).
__device__ __forceinline__ double rsqrt_nr(double x)
{
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  y = fma(y, fma(-hx * y, y, 0.5), y);
  y = fma(y, fma(-hx * y, y, 0.5), y);
  return y;
}

// J1(x), x >= 0 (role of gsl_sf_bessel_J1 in src/UpcCrossSection.cpp:189).
// J1 on its two branches (x <= 8: Chebyshev series in (x/4)^2; x > 8:
// modulus/phase form), callable separately so a caller can evaluate the
// lanes of one branch together.
__device__ __forceinline__ double bessel_j1_small(double x)
{
  const double u = x * x * (1.0 / 32.0) - 1.0;
  return x * clenshaw(kJ1Small, u);
}

__device__ __forceinline__ double bessel_j1_large(double x)
{
  constexpr double kInvSqrtPi = 0.56418958354775628695;
  const double y = rsqrt_nr(x);
  const double r = y * y;  // 1/x
  const double v = fma(128.0 * r, r, -1.0);
  double P = kJ1P[kJ1P_N - 1], Q = kJ1Q[kJ1Q_N - 1];
#pragma unroll
  for (int k = kJ1P_N - 2; k >= 0; --k) {
    P = fma3(v, P, kJ1P[k]);
    Q = fma3(v, Q, kJ1Q[k]);
  }
  Q *= 8.0 * r;
  double s, c;
  sincos_medium(x, &s, &c);
  return (P * (s - c) + Q * (s + c)) * (y * kInvSqrtPi);
}

__device__ __forceinline__ double bessel_j1(double x)
{
  return x <= 8.0 ? bessel_j1_small(x) : bessel_j1_large(x);
}

// K0(x), K1(x) for x > 0 (role of gsl_sf_bessel_K0/K1 in
// src/UpcCrossSection.cpp:171-172).  Underflows quietly to 0 at large x.
__device__ __forceinline__ void bessel_k01(double x, double* k0, double* k1)
{
  if (x <= 1.0) {
    const double y = 0.25 * x * x;
    const double lx = log(0.5 * x);
    const double i0 = horner(kI0Ser, y);
    const double s0 = horner(kS0Ser, y);
    const double i1s = horner(kI1Ser, y);
    const double s1 = horner(kS1Ser, y);
    constexpr double kEuler = 0.57721566490153286060651209008240243;
    *k0 = -(lx + kEuler) * i0 + s0;
    *k1 = 1.0 / x + lx * (0.5 * x * i1s) - 0.25 * x * s1;
    return;
  }
  const double t = 1.0 / x;
  double s0, s1;
  if (x <= 4.0) {
    const double u = (2.0 * t - 1.25) * (1.0 / 0.75);
    clenshaw2(kK0A, kK1A, u, &s0, &s1);
  } else {
    const double u = 8.0 * t - 1.0;
    clenshaw2(kK0B, kK1B, u, &s0, &s1);
  }
  const double e = exp(-x) * sqrt(t);
  *k0 = s0 * e;
  *k1 = s1 * e;
}

// e^x K0(x), e^x K1(x): the same series, without the exponential underflow.
__device__ __forceinline__ void bessel_k01_scaled(double x, double* k0s, double* k1s)
{
  if (x <= 1.0) {
    double k0, k1;
    bessel_k01(x, &k0, &k1);
    const double e = exp(x);
    *k0s = k0 * e;
    *k1s = k1 * e;
    return;
  }
  const double t = 1.0 / x;
  double s0, s1;
  if (x <= 4.0) {
    const double u = (2.0 * t - 1.25) * (1.0 / 0.75);
    clenshaw2(kK0A, kK1A, u, &s0, &s1);
  } else {
    const double u = 8.0 * t - 1.0;
    clenshaw2(kK0B, kK1B, u, &s0, &s1);
  }
  const double e = sqrt(t);
  *k0s = s0 * e;
  *k1s = s1 * e;
}

// Abramowitz & Stegun 9.8.3/9.8.4: the polynomial I1 of TMath::BesselI1.
__device__ __forceinline__ double bessel_i1_as(double x)
{
  const double ax = fabs(x);
  if (ax < 3.75) {
    const double t = x / 3.75;
    const double y = t * t;
    return x * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 +
               y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  }
  const double y = 3.75 / ax;
  const double r = (exp(ax) / sqrt(ax)) *
                   (0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 + y * (0.163801e-2 +
                    y * (-0.1031555e-1 + y * (0.2282967e-1 + y * (-0.2895312e-1 +
                    y * (0.1787654e-1 + y * -0.420059e-2))))))));
  return x < 0 ? -r : r;
}

// Abramowitz & Stegun 9.8.7/9.8.8: the polynomial K1 of TMath::BesselK1, used
// by the Starlight breakup sum (src/UpcCrossSection.cpp:982-999).
__device__ __forceinline__ double bessel_k1_as(double x)
{
  if (x <= 0) return 0.;
  if (x <= 2) {
    const double y = x * x / 4;
    return (log(x / 2) * bessel_i1_as(x)) +
           (1. / x) * (1. + y * (0.15443144 + y * (-0.67278579 + y * (-0.18156897 +
            y * (-0.1919402e-1 + y * (-0.110404e-2 + y * (-0.4686e-4)))))));
  }
  const double y = 2 / x;
  return (exp(-x) / sqrt(x)) *
         (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 + y * (0.1504268e-1 +
          y * (-0.780353e-2 + y * (0.325614e-2 + y * (-0.68245e-3)))))));
}

// LDS written by some lanes of a wavefront, then read by others of the same
// wavefront: order the accesses without a workgroup barrier.
__device__ __forceinline__ void wave_sync_lds()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Natural cubic spline on knots x_i = x0 + i*dx (each op rounded separately,
// as the host loop that built the knots did).  coef[i] = (y_i, b_i, c_i, d_i)
// for interval i; the interval is chosen as gsl_interp_bsearch would:
// x_i <= v < x_{i+1}, i in [0, n-2].
// ---------------------------------------------------------------------------
__device__ __forceinline__ double knot(double x0, double dx, int i)
{
#pragma clang fp contract(off)
  return x0 + (double)i * dx;
}

// The estimate (v - x0) / dx and the knot positions i dx each carry a
// rounding error below 1e-9 of a knot spacing for the tables here (dx >=
// 1e-6 with |x| <= 1e3), so the estimated interval is off by at most one:
// one conditional step down, then one up, gives x_i <= v < x_{i+1}.
__device__ __forceinline__ int spline_interval(double x0, double dx, double inv_dx, int n, double v,
                                               double* xi)
{
  int i = (int)((v - x0) * inv_dx);
  i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  double xl = knot(x0, dx, i);
  const bool down = i > 0 && v < xl;
  const double xd = knot(x0, dx, i - 1);
  i = down ? i - 1 : i;
  xl = down ? xd : xl;
  const double xh = knot(x0, dx, i + 1);
  const bool up = i < n - 2 && !(v < xh);
  i = up ? i + 1 : i;
  *xi = up ? xh : xl;
  return i;
}

__device__ __forceinline__ double spline_eval(const double4* __restrict__ coef, double x0, double dx,
                                              double inv_dx, int n, double v)
{
  double xl;
  const int i = spline_interval(x0, dx, inv_dx, n, v, &xl);
  const double4 c = coef[i];
  const double d = v - xl;
  return c.x + d * (c.y + d * (c.z + d * c.w));
}

// The same cubic on the estimated interval alone (no one-knot correction).
// When v lies within rounding of a knot the estimate may name the
// neighbouring interval; the spline is C2, so that cubic agrees at v to a
// few ulp.  For sums of many spline values, where the exact interval choice
// does not matter.
__device__ __forceinline__ double spline_eval_near(const double4* __restrict__ coef, double x0, double dx,
                                                   double inv_dx, int n, double v)
{
  int i = (int)((v - x0) * inv_dx);
  i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  const double4 c = coef[i];
  const double d = v - knot(x0, dx, i);
  return c.x + d * (c.y + d * (c.z + d * c.w));
}

__device__ __forceinline__ double spline_deriv(const double4* __restrict__ coef, double x0, double dx,
                                               double inv_dx, int n, double v)
{
  double xl;
  const int i = spline_interval(x0, dx, inv_dx, n, v, &xl);
  const double4 c = coef[i];
  const double d = v - xl;
  return c.y + d * (2.0 * c.z + 3.0 * d * c.w);
}

// ---------------------------------------------------------------------------
// QUADPACK dqagse building blocks (QK21, dqpsrt, dqelg)
// ---------------------------------------------------------------------------
__device__ constexpr double kXgk[11] = {
    0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
    0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
    0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
    0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
    0.294392862701460198131126603103866, 0.148874338981631210884826001129720,
    0.000000000000000000000000000000000};
__device__ constexpr double kWgk[11] = {
    0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
    0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
    0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
    0.123491976262065851077208980748500, 0.134709217311473325928054001771707,
    0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
    0.149445554002916905664936468389821};
__device__ constexpr double kWg[5] = {
    0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
    0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
    0.295524224714752870173892994651338};

struct QK21Out {
  double result, abserr, resabs, resasc;
};

// dqk21 with the dqagse error rescaling; same summation order as QUADPACK.
template <class F>
__device__ __forceinline__ QK21Out qk21(const F& f, double a, double b)
{
  constexpr double epmach = DBL_EPSILON, uflow = DBL_MIN;
  double fv1[10], fv2[10];
  const double centr = 0.5 * (a + b);
  const double hlgth = 0.5 * (b - a);
  const double dhlgth = fabs(hlgth);
  double resg = 0.0;
  const double fc = f(centr);
  double resk = kWgk[10] * fc;
  double rabs = fabs(resk);
#pragma unroll
  for (int j = 1; j <= 5; ++j) {
    const int jtw = 2 * j - 1; // 0-based index of the Fortran xgk(2j)
    const double absc = hlgth * kXgk[jtw];
    const double f1 = f(centr - absc);
    const double f2 = f(centr + absc);
    fv1[jtw] = f1;
    fv2[jtw] = f2;
    const double fsum = f1 + f2;
    resg = resg + kWg[j - 1] * fsum;
    resk = resk + kWgk[jtw] * fsum;
    rabs = rabs + kWgk[jtw] * (fabs(f1) + fabs(f2));
  }
#pragma unroll
  for (int j = 1; j <= 5; ++j) {
    const int jtwm1 = 2 * j - 2; // xgk(2j-1)
    const double absc = hlgth * kXgk[jtwm1];
    const double f1 = f(centr - absc);
    const double f2 = f(centr + absc);
    fv1[jtwm1] = f1;
    fv2[jtwm1] = f2;
    const double fsum = f1 + f2;
    resk = resk + kWgk[jtwm1] * fsum;
    rabs = rabs + kWgk[jtwm1] * (fabs(f1) + fabs(f2));
  }
  const double reskh = resk * 0.5;
  double rasc = kWgk[10] * fabs(fc - reskh);
#pragma unroll
  for (int j = 0; j < 10; ++j) rasc = rasc + kWgk[j] * (fabs(fv1[j] - reskh) + fabs(fv2[j] - reskh));
  QK21Out o;
  o.result = resk * hlgth;
  rabs = rabs * dhlgth;
  rasc = rasc * dhlgth;
  double err = fabs((resk - resg) * hlgth);
  if (rasc != 0.0 && err != 0.0) {
    const double t15 = 200.0 * err / rasc;
    const double sc = t15 * sqrt(t15);
    err = rasc * (sc < 1.0 ? sc : 1.0);
  }
  if (rabs > uflow / (50.0 * epmach)) {
    const double m = (epmach * 50.0) * rabs;
    if (m > err) err = m;
  }
  o.abserr = err;
  o.resabs = rabs;
  o.resasc = rasc;
  return o;
}

// The 21 nodes of a QK21 pass in the order the passes below visit them:
// e = 0 the centre, e = 1..10 the Gauss abscissae xgk(2j) in (left, right)
// pairs, e = 11..20 the Kronrod-only abscissae xgk(2j - 1) in pairs.
__device__ __forceinline__ double qk21_abscissa(int e, double centr, double hlgth)
{
  const bool gauss = e <= 10;
  const int p = gauss ? (e - 1) >> 1 : (e - 11) >> 1;
  const int side = gauss ? (e - 1) & 1 : (e - 11) & 1;
  const int xi = e == 0 ? 10 : (gauss ? 2 * p + 1 : 2 * p);
  const double absc = hlgth * kXgk[xi];
  return e == 0 ? centr : (side ? centr + absc : centr - absc);
}

// dqk21 as a rolled loop over the 21 nodes (e = 0 the centre, e = 1..10
// the Gauss pairs, e = 11..20 the Kronrod-only pairs; qk21_abscissa): the
// same operations in the same order as qk21() above.  fx(e, x) gives the
// integrand at node e, abscissa x; gx(j, side) gives back the value of
// pair j (kXgk index), side 0 (centr - absc) or 1 (centr + absc), for the
// mean deviation once the Kronrod mean is known.
template <class GX>
__device__ __forceinline__ QK21Out qk21_tail(const GX& gx, double resg, double resk, double rabs, double fc,
                                             double hlgth);

template <class FX, class GX>
__device__ __forceinline__ QK21Out qk21_core(const FX& fx, const GX& gx, double a, double b)
{
  const double centr = 0.5 * (a + b);
  const double hlgth = 0.5 * (b - a);
  double resg = 0.0, resk = 0.0, rabs = 0.0, fc = 0.0, f1 = 0.0;
#pragma unroll 1
  for (int e = 0; e < 21; ++e) {
    const bool gauss = e <= 10;
    const int p = gauss ? (e - 1) >> 1 : (e - 11) >> 1;
    const int side = gauss ? (e - 1) & 1 : (e - 11) & 1;
    const int xi = e == 0 ? 10 : (gauss ? 2 * p + 1 : 2 * p);
    const double fx_e = fx(e, qk21_abscissa(e, centr, hlgth));
    if (e == 0) {
      fc = fx_e;
      resk = kWgk[10] * fc;
      rabs = fabs(resk);
    } else if (side == 0) {
      f1 = fx_e;
    } else {
      const double f2 = fx_e;
      const double fsum = f1 + f2;
      if (gauss) resg = resg + kWg[p] * fsum;
      resk = resk + kWgk[xi] * fsum;
      rabs = rabs + kWgk[xi] * (fabs(f1) + fabs(f2));
    }
  }
  return qk21_tail(gx, resg, resk, rabs, fc, hlgth);
}

// The end of dqk21 once the node sums are known: the mean deviation (gx as
// in qk21_core) and the error estimate.
template <class GX>
__device__ __forceinline__ QK21Out qk21_tail(const GX& gx, double resg, double resk, double rabs, double fc,
                                             double hlgth)
{
  constexpr double epmach = DBL_EPSILON, uflow = DBL_MIN;
  const double dhlgth = fabs(hlgth);
  const double reskh = resk * 0.5;
  double rasc = kWgk[10] * fabs(fc - reskh);
#pragma unroll
  for (int j = 0; j < 10; ++j) rasc = rasc + kWgk[j] * (fabs(gx(j, 0) - reskh) + fabs(gx(j, 1) - reskh));
  QK21Out o;
  o.result = resk * hlgth;
  rabs = rabs * dhlgth;
  rasc = rasc * dhlgth;
  double err = fabs((resk - resg) * hlgth);
  if (rasc != 0.0 && err != 0.0) {
    const double t15 = 200.0 * err / rasc;
    const double sc = t15 * sqrt(t15);
    err = rasc * (sc < 1.0 ? sc : 1.0);
  }
  if (rabs > uflow / (50.0 * epmach)) {
    const double m = (epmach * 50.0) * rabs;
    if (m > err) err = m;
  }
  o.abserr = err;
  o.resabs = rabs;
  o.resasc = rasc;
  return o;
}

// Node e of pair j (kXgk index), side 0 or 1 (the inverse of the mapping in
// qk21_core).
__device__ __forceinline__ int qk21_node(int j, int side) { return (j & 1) ? j + side : 11 + j + side; }

// qk21_core evaluating the integrand: the 20 off-centre values are parked
// at fv[(2 j + side) * stride] until the mean deviation needs them.
template <class FX>
__device__ __forceinline__ QK21Out qk21_sum(const FX& fx, double a, double b, double* fv, int stride)
{
  return qk21_core(
      [&](int e, double x) {
        const double v = fx(e, x);
        if (e != 0) {
          const int gauss = e <= 10;
          const int p = gauss ? (e - 1) >> 1 : (e - 11) >> 1;
          const int side = gauss ? (e - 1) & 1 : (e - 11) & 1;
          const int xi = gauss ? 2 * p + 1 : 2 * p;
          fv[(2 * xi + side) * stride] = v;
        }
        return v;
      },
      [&](int j, int side) { return fv[(2 * j + side) * stride]; }, a, b);
}

// qk21_core over values already evaluated (node e's value at v[e * stride]),
// as two loops over pairs: the Gauss pairs (nodes 1..10), then the
// Kronrod-only pairs (11..20), summed in qk21_core's order.  (qk21_core's
// loop over the 21 nodes branches on each node's kind and waits for a scalar
// load of its weights on every trip.)  The Gauss loop is unrolled (its ten
// loads issued together); unrolling the second as well costs the QAGS kernel
// a spilled VGPR.
__device__ __forceinline__ QK21Out qk21_stored(const double* v, int stride, double a, double b)
{
  const double hlgth = 0.5 * (b - a);
  const double fc = v[0];
  double resg = 0.0, resk = kWgk[10] * fc;
  double rabs = fabs(resk);
#pragma unroll
  for (int p = 0; p < 5; ++p) {
    const double f1 = v[(2 * p + 1) * stride], f2 = v[(2 * p + 2) * stride];
    const double fsum = f1 + f2;
    resg = resg + kWg[p] * fsum;
    resk = resk + kWgk[2 * p + 1] * fsum;
    rabs = rabs + kWgk[2 * p + 1] * (fabs(f1) + fabs(f2));
  }
#pragma unroll 1
  for (int p = 0; p < 5; ++p) {
    const double f1 = v[(11 + 2 * p) * stride], f2 = v[(12 + 2 * p) * stride];
    const double fsum = f1 + f2;
    resk = resk + kWgk[2 * p] * fsum;
    rabs = rabs + kWgk[2 * p] * (fabs(f1) + fabs(f2));
  }
  return qk21_tail([&](int j, int side) { return v[qk21_node(j, side) * stride]; }, resg, resk, rabs, fc, hlgth);
}

// dqk21 of f on [a, b] through qk21_sum.
template <class F>
__device__ __forceinline__ QK21Out qk21_lds(const F& f, double a, double b, double* fv, int stride)
{
  return qk21_sum([&](int, double x) { return f(x); }, a, b, fv, stride);
}

// dqelg (Wynn epsilon) on a 1-based table of length >= 55.
__device__ __forceinline__ void qelg(int* n, double* epstab, double* result, double* abserr, double* res3la,
                            int* nres)
{
  constexpr double epmach = DBL_EPSILON, oflow = DBL_MAX;
  constexpr int limexp = 50;
  int i, ib, ie, ib2, indx, k1, k2, k3, newelm, num;
  double delta1, delta2, delta3, e0, e1, e1abs, e2, e3, err1, err2, err3, error, res, ss, epsinf,
      tol1, tol2, tol3;
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

}  // namespace upc
