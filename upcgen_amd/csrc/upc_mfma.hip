// MFMA form of the two-photon luminosity grid.
//
// Reference: calcTwoPhotonLumi / calcTwoPhotonLumiPol,
// src/UpcCrossSection.cpp:221-335 (the same sum the FP64 form evaluates node
// by node).
//
// The flux rows are the FP64 rows of the FP64 form (point-charge flux and the
// QAGS form-factor flux).  The reference's QAGS runs at epsabs = epsrel =
// 1e-4, and where the flux is small its result is mostly quadrature error,
// which dominates L in the large-k corner bins; so the rows themselves are
// never approximated.  What is recast is the b1/b2/phi kernel: with
//   W_i   = N(b_i, k) b_i (b_h - b_l)                (one flux row)
//   g(b)  = P(min(b, 20)) G(b),  G = 1 for b >= 20
//   K(b1, b2) = sum_k w_k g(|b1 + b2 e^{i phi_k}|),   D = K - S0
// the node sum is L = 2 pi^2 M [S0 (sum W1)(sum W2) + sum_ij W1_i W2_j D_ij]
// and D is interpolated on one uniform grid B_u = u h shared by every row:
//   What[r][u] = sum_i W_i l_u(b_i)     (4-point Lagrange weights)
//   L = 2 pi^2 M [S0 (sum W1)(sum W2) + What1^T D What2],  D_uv = D(B_u, B_v).
// At b = 20 fm g is not smooth: P is clamped there (slope jump) and G_AA
// switches from its spline to exactly 1 (value jump J = g(20) - g(20-)).
// Interpolating that edge costs ~1e-5 of L at h = 0.2 fm, so it is split off:
//   g = g_c + e,   e(b) = [b < 20] (-J + s (400 - b^2)),  s = -g'(20-) / 40,
// which leaves g_c continuous with a continuous slope at 20 fm (e jumps by J
// and has slope -40 s = g'(20-) just below 20).  Only D_c (from g_c) is
// interpolated; the ramp is a quadratic in (b1, b2), which the 4-point
// Lagrange interpolation reproduces exactly.  The edge part
//   sum_ij W1_i W2_j sum_k w_k e(b_k(i, j))
// is evaluated exactly at the rows' own node pairs (kernel k_mfma_edge): for
// fixed (i, k) the j with b_k < 20 form one contiguous range (b_k^2 is a
// parabola in b2), and with b_k^2 = b1^2 + b2^2 + 2 b1 b2 c the sum over
// that range is a difference of prefix sums of W2, W2 b2 and W2 b2^2.
// D_c is symmetric, vanishes for |B_u - B_v| >= 20 fm (every node is then at or
// beyond 20 fm, where K = S0 in the same arithmetic) and for B >= 20 /
// sqrt(1 - c^2) with c the extreme node cosine, whatever the partner.  So the
// contraction is a banded GEMM  Y = What2 D  over all points at once, fused
// with the dot  Y . What1.  It runs on MFMA as split-bf16 (x = hi + lo,
// hi*hi + hi*lo + lo*hi, FP32 accumulation) or as FP32 MFMA; the dot and the
// S0 term are FP64.  Each What row is scaled by a power of two (kept in FP64)
// that brings its values below 1, so the FP32 range never limits a row.
#include "upc_flux.h"

#include <cmath>
#include <cstdint>

namespace upc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int kMfmaWaves = 4;      // waves per block in the contraction kernel
constexpr int kMfmaTileGroup = 4;  // column tiles per A-fragment load of a single grid

// e(b) of the edge split below b = 20 (x = b^2 < 400); J = edge[0],
// -g'(20-) = edge[1], so s = -g'(20-) / 40 = edge[1] / 40.
__device__ __forceinline__ double edge_e(double x, double J, double sl)
{
  return -J + sl * (1.0 / 40.0) * (400.0 - x);
}

// J and -g'(20-) of g(b) = P(b) G(b) (P = 1 in mode 1) from the splines,
// and the five phi-node cosines (edge[2..6]).
__global__ void k_mfma_edge_params(int mode, const double4* __restrict__ gaa, const double4* __restrict__ bk,
                                   const DevScalars* __restrict__ sc, double* __restrict__ edge)
{
  const double db = 20. / (kNB - 1);
  const double G = spline_eval(gaa, 0.0, db, 1.0 / db, kNB, 20.0);
  const double dG = spline_deriv(gaa, 0.0, db, 1.0 / db, kNB, 20.0);
  double P = 1., dP = 0.;
  if (mode != 1) {
    P = sc->p20;
    dP = spline_deriv(bk, kBkMin, kDBk, 1.0 / kDBk, kNBkUsed, 20.0);
  }
  edge[0] = P - P * G;  // g(20) = P(20) exactly; g(20-) = P(20) G(20)
  edge[1] = -(dP * G + P * dG);
#pragma unroll
  for (int k = 0; k < 5; ++k) edge[2 + k] = cos(kPi * kGLx5[k]);  // the phi-node cosines, once
}

struct MfmaGrid {
  double h;       // node spacing of the shared grid (fm)
  double inv_h;
  double bstar;   // D(b, .) = 0 for b >= bstar
  int nu;         // nodes in use
  int nu_pad;     // row length (multiple of 32)
  int band;       // D_uv = 0 for |u - v| >= band
  int npol;       // 1, or 2 for the polarised pair (D_s, D_p)
};

// Cancellation guard.  A bin's value is s0 sum(w1) sum(w2) (FP64) plus the
// MFMA product W1 D W2^T (D = the phi sum minus its far-field value s0, in
// [-s0, 0] below 20 fm) plus the exact edge part.  The product's error is
// a fraction of order 1e-7 of s0 sum(w1) sum(w2) (split-bf16 or FP32
// products, FP32 sums), so where the product cancels most of the far-field
// term (flux concentrated below 2R, where G_AA and the breakup factor vanish:
// high M at low sqrt(s)) the bin's relative error grows with the ratio
// s0 sum(w1) sum(w2) / |L| (about 2.5e-6 times it, measured on Au-Au at
// 200 GeV).  Work items whose ratio exceeds the limit are recomputed by the
// FP64 contraction from the same flux rows; the default keeps the bins of
// every tested system within 1e-5 (none is flagged on Pb-Pb, Xe-Xe or O-O
// at LHC energies).
constexpr double kMfmaCancelLimit = 2.0;

struct MfmaState {
  double h = 0.2;
  int prec = UPC_FORM_MFMA;
  MfmaGrid g{};
  size_t d_cap = 0, rows_cap = 0;
  uint16_t* d_Dh = nullptr;   // [npol][nu_pad][nu_pad] bf16 hi
  uint16_t* d_Dl = nullptr;   // bf16 lo
  float* d_D32 = nullptr;     // FP32 variant
  uint16_t* d_Wh = nullptr;   // [rows][nu_pad]
  uint16_t* d_Wl = nullptr;
  float* d_W32 = nullptr;
  double* d_wscale = nullptr; // [rows]
  double* d_wsum = nullptr;   // [rows] sum of all 120 W_i
  int* d_ulast = nullptr;     // [rows] last nonzero What column
  double* d_edge_sum = nullptr;  // [npol][points] exact edge part
  double* d_edge = nullptr;      // J, -g'(20-), the five phi-node cosines
  unsigned char* d_need = nullptr;  // [work items] 1: recomputed by the FP64 contraction
  unsigned long long* d_ksteps = nullptr;  // [0] MFMA tile steps, [1] work items recomputed in FP64
  unsigned long long last[2] = {0, 0};     // host copy of d_ksteps after the last single-context grid
  double cancel_limit = kMfmaCancelLimit;
};

static MfmaState* state(upc_ctx* c)
{
  if (!c->mfma) c->mfma = new MfmaState();
  return static_cast<MfmaState*>(c->mfma);
}

__device__ __forceinline__ uint16_t to_bf16(float f)
{
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite input)
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ void split_store(double x, uint16_t* hi, uint16_t* lo, float* f32, long i)
{
  if (f32) {
    f32[i] = (float)x;
  } else {
    const uint16_t h = to_bf16((float)x);
    hi[i] = h;
    lo[i] = to_bf16((float)(x - (double)from_bf16(h)));
  }
}

// ---------------------------------------------------------------------------
// D on the shared grid
// ---------------------------------------------------------------------------
template <bool POL>
__global__ __launch_bounds__(256) void k_mfma_kernel_matrix(MfmaGrid g, int mode, const double4* __restrict__ gaa,
                                                             const double4* __restrict__ bk,
                                                             const DevScalars* __restrict__ sc,
                                                             const double* __restrict__ edge,
                                                             uint16_t* __restrict__ Dh, uint16_t* __restrict__ Dl,
                                                             float* __restrict__ D32)
{
  const long n = (long)g.nu_pad * g.nu_pad;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int u = (int)(idx / g.nu_pad), v = (int)(idx - (long)u * g.nu_pad);
  double cosk[5], w5[10];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double c = cos(kPi * kGLx5[k]), sn = sin(kPi * kGLx5[k]);
    cosk[k] = c;
    w5[k] = kGLw5[k];
    w5[5 + k] = sn * sn;
  }
  const double p20 = sc->p20;
  double ds = 0, dp = 0;
  if (u < g.nu && v < g.nu && abs(u - v) < g.band) {
    const double br = mode != 1 ? p20 : 1.;
    double s0s = 0, s0p = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (POL) {
        s0s += br * 1. * w5[k] * cosk[k] * cosk[k];
        s0p += br * 1. * w5[k] * w5[5 + k];
      } else {
        s0s += br * 1. * w5[k];
      }
    }
    const double db = 20. / (kNB - 1);
    const double J = edge[0], sl = edge[1];
    const double b1 = u * g.h, b2 = v * g.h;
    double ks = 0, kp = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double c = cosk[k];
      const double x = node_b2(b1, b2, POL ? -c : c);
      const double b = sqrt(x);
      double gc = br;
      if (b < 20.) {
        const double breakup = mode != 1 ? spline_eval(bk, kBkMin, kDBk, 1.0 / kDBk, kNBkUsed, b) : 1.;
        gc = breakup * spline_eval(gaa, 0.0, db, 1.0 / db, kNB, b) - edge_e(x, J, sl);
      }
      if (POL) {
        ks += gc * w5[k] * c * c;
        kp += gc * w5[k] * w5[5 + k];
      } else {
        ks += gc * w5[k];
      }
    }
    ds = ks - s0s;
    dp = kp - s0p;
  }
  split_store(ds, Dh, Dl, D32, idx);
  if (POL) split_store(dp, Dh ? Dh + n : nullptr, Dl ? Dl + n : nullptr, D32 ? D32 + n : nullptr, idx);
}

// ---------------------------------------------------------------------------
// Projection of the flux rows onto the shared grid: one wave per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mfma_project(int nrows, const double* __restrict__ row_logd,
                                                      const double* __restrict__ flux, FluxParams fp, MfmaGrid g,
                                                      uint16_t* __restrict__ Wh, uint16_t* __restrict__ Wl,
                                                      float* __restrict__ W32, double* __restrict__ wscale,
                                                      double* __restrict__ wsum, int* __restrict__ ulast)
{
  __shared__ double sW[4][kNB1], sL[4][kNB1][4];
  __shared__ int sI[4][kNB1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = blockIdx.x * 4 + wv;
  const bool live = r0 < nrows;
  const int r = live ? r0 : nrows - 1;
  const double logd = row_logd[r];
  LogGrid lg;
  lg.init(fp.bmin, logd);
  double tot = 0, tabs = 0;
  int ulim = -1;  // last column any node feeds: all later columns are 0
  for (int j = lane; j < kNB1; j += 64) {
    double bl, bh;
    const double b = row_node(fp.bmin, logd, j, &bl, &bh);
    const double w = flux[(long)r * kNB1 + j] * b * (bh - bl);
    tot += w;
    tabs += fabs(w);
    sW[wv][j] = w;
    int i0 = INT_MAX;
    if (b < g.bstar) {
      const double t = b * g.inv_h;
      i0 = (int)floor(t) - 1;
      i0 = i0 < 0 ? 0 : (i0 > g.nu - 4 ? g.nu - 4 : i0);
      const double x = t - i0;  // nodes at x = 0, 1, 2, 3
      sL[wv][j][0] = -(x - 1.) * (x - 2.) * (x - 3.) / 6.;
      sL[wv][j][1] = x * (x - 2.) * (x - 3.) / 2.;
      sL[wv][j][2] = -x * (x - 1.) * (x - 3.) / 2.;
      sL[wv][j][3] = x * (x - 1.) * (x - 2.) / 6.;
    }
    sI[wv][j] = i0;
    ulim = i0 == INT_MAX ? ulim : max(ulim, i0 + 3);
  }
  for (int o = 32; o > 0; o >>= 1) {
    tot += __shfl_xor(tot, o);
    tabs += __shfl_xor(tabs, o);
    ulim = max(ulim, __shfl_xor(ulim, o));
  }
  __syncthreads();
  // gather: node i feeds columns i0_i .. i0_i + 3; i0 is nondecreasing in i.
  auto column = [&](int u) {
    // first i with i0_i >= u - 3: i0_i >= u - 3 when b_i >= (u - 2) h (the
    // clamps of i0 aside), guessed on the log grid and stepped to exactness
    int lo = u <= 3 ? 0 : lg.first_ge((u - 2) * g.h, kNB1);
    while (lo > 0 && sI[wv][lo - 1] >= u - 3) --lo;
    while (lo < kNB1 && sI[wv][lo] < u - 3) ++lo;
    double a = 0;
    for (int i = lo; i < kNB1 && sI[wv][i] <= u; ++i) a += sW[wv][i] * sL[wv][i][u - sI[wv][i]];
    return a;
  };
  // One pass: each column is scaled by a power of two 2^-e with 4 sum|W| <
  // 2^e, so every stored value is below 1 in magnitude (|column u| <= sum_i
  // |W_i| |l_i(u)|, and the 4-point weights stay below 4 even where i0 is
  // clamped at the grid ends) and the FP32 range never limits a row.  The
  // scaling is exact, so the bf16 / FP32 split keeps the row's own mantissas.
  const int e = tabs > 0 ? ilogb(4.0 * tabs) + 1 : 0;
  const double sc = ldexp(1.0, -e);
  const int uend = min(g.nu, ulim + 1);
  int last = -1;
  for (int u = lane; u < g.nu_pad; u += 64) {
    const double a = u < uend ? column(u) : 0.0;
    if (a != 0) last = u;
    if (live) split_store(a * sc, Wh, Wl, W32, (long)r * g.nu_pad + u);
  }
  for (int o = 32; o > 0; o >>= 1) last = max(last, __shfl_xor(last, o));
  if (live && lane == 0) {
    wscale[r] = ldexp(1.0, e);
    wsum[r] = tot;
    ulast[r] = last;
  }
}

// ---------------------------------------------------------------------------
// Edge part: sum_i W1_i sum_k w_k t_k sum_{j: b_k(i,j) < 20} W2_j e(b_k(i,j)),
// t_k = 1 (unpolarised) or cos^2 / sin^2 (polarised), at the rows' own nodes.
// One wave per point; lanes take (i, k) pairs.
// ---------------------------------------------------------------------------
template <bool POL>
__global__ __launch_bounds__(256) void k_mfma_edge(int npts, int mirror_out, const double* __restrict__ row_logd,
                                                   const double* __restrict__ flux, FluxParams fp,
                                                   const double* __restrict__ edge, double* __restrict__ out)
{
  // spre[.][m][j] = sum_{j' < j} W2_j' b2_j'^m, m = 0, 1, 2
  __shared__ double sb1[4][kNB1], sb2[4][kNB2], sw1[4][kNB1], spre[4][3][kNB2 + 1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwork = lumi_work_items(npts, fp.mirror_ny, mirror_out);
  const int q0 = blockIdx.x * 4 + wv;
  const bool live = q0 < nwork;
  int p, pm;  // the edge term of a mirror bin equals its partner's; only p is stored
  work_point(live ? q0 : nwork - 1, fp.mirror_ny, mirror_out, &p, &pm);
  const int r2 = k2_row(p, npts, fp.mirror_ny);
  const double logd1 = row_logd[p], logd2 = row_logd[r2];
  LogGrid lg2;
  lg2.init(fp.bmin, logd2);
  double w2[2] = {0.0, 0.0}, b2v[2] = {0.0, 0.0};  // nodes lane and lane + 64 of row r2
  for (int h = 0; h < 2; ++h) {
    const int j = lane + 64 * h;
    if (j >= kNB1) break;
    double bl, bh;
    const double b1 = row_node(fp.bmin, logd1, j, &bl, &bh);
    sb1[wv][j] = b1;
    sw1[wv][j] = flux[(long)p * kNB1 + j] * b1 * (bh - bl);
    const double b2 = row_node(fp.bmin, logd2, j, &bl, &bh);
    sb2[wv][j] = b2;
    b2v[h] = b2;
    w2[h] = flux[(long)r2 * kNB1 + j] * b2 * (bh - bl);
  }
  // prefix sums of W2, W2 b2, W2 b2^2: wave scans of the two halves
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double c0 = w2[0] * (m == 0 ? 1.0 : (m == 1 ? b2v[0] : b2v[0] * b2v[0]));
    double c1 = w2[1] * (m == 0 ? 1.0 : (m == 1 ? b2v[1] : b2v[1] * b2v[1]));
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const double u0 = __shfl_up(c0, o), u1 = __shfl_up(c1, o);
      if (lane >= o) {
        c0 += u0;
        c1 += u1;
      }
    }
    const double half = __shfl(c0, 63);
    spre[wv][m][lane + 1] = c0;
    if (lane + 64 < kNB2) spre[wv][m][lane + 65] = half + c1;
    if (lane == 0) spre[wv][m][0] = 0;
  }
  __syncthreads();
  if (!live) return;
  const double J = edge[0], s = edge[1] * (1.0 / 40.0);  // e = -J + s (400 - b^2) below 20
  double cosk[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) cosk[k] = edge[2 + k];
  double acc_s = 0, acc_p = 0;
  for (int t = lane; t < kNB1 * 5; t += 64) {
    const int i = t / 5, k = t - 5 * (t / 5);
    double cs = cosk[0];
#pragma unroll
    for (int q = 1; q < 5; ++q) cs = k == q ? cosk[q] : cs;
    const double c = POL ? -cs : cs;
    const double b1 = sb1[wv][i];
    // no node pair reaches b < 20 when even the closest approach b1 sin(phi) does not
    if (b1 * b1 * (1.0 - c * c) >= 400.0 * (1.0 + 1e-12)) continue;
    auto bk2 = [&](int j) { return node_b2(b1, sb2[wv][j], c); };
    // sqrt(x) < 20 exactly as the node loop tests it: x >= 400 is never
    // inside, and only x within rounding of 400 needs the square root
    auto inside = [&](double x) { return x < 400. && (x < 399.99999 || sqrt(x) < 20.); };
    // b_k^2 is smallest at b2 = -b1 c: the last node below it or the first above
    // (first j with b2_j >= vtx: a log-grid guess stepped to exactness)
    const double vtx = -b1 * c;
    int lo = lg2.first_ge(vtx, kNB2);
    while (lo > 0 && sb2[wv][lo - 1] >= vtx) --lo;
    while (lo < kNB2 && sb2[wv][lo] < vtx) ++lo;
    int jc = -1;
    const double xlo = lo < kNB2 ? bk2(lo) : 1e300, xlm = lo > 0 ? bk2(lo - 1) : 1e300;
    if (inside(xlo)) jc = lo;
    if (inside(xlm) && (jc < 0 || xlm < xlo)) jc = lo - 1;
    if (jc < 0) continue;
    // The pass region of b_k^2 < X around jc ends near the roots
    // b2 = -b1 c -+ sqrt(X - b1^2 (1 - c^2)): each end is guessed on the log
    // grid and stepped with pass() itself (false ... true on [0, jc], true ...
    // false on [jc, n)), so the ends are those a bisection would find.
    auto root = [&](double X, double sgn) {
      const double d = X - b1 * b1 * (1.0 - c * c);
      return vtx + sgn * sqrt(d > 0.0 ? d : 0.0);
    };
    // first node in [0, jc] that passes (b_k falls towards jc)
    auto first_in = [&](auto pass, double X) {
      int a = lg2.first_ge(root(X, -1.0), kNB2);
      a = a > jc ? jc : a;
      while (a > 0 && pass(bk2(a - 1))) --a;
      while (a < jc && !pass(bk2(a))) ++a;
      return a;
    };
    // last node in [jc, n) that passes (b_k rises after jc)
    auto last_in = [&](auto pass, double X) {
      int a = lg2.first_ge(root(X, 1.0), kNB2) - 1;
      a = a < jc ? jc : (a > kNB2 - 1 ? kNB2 - 1 : a);
      while (a < kNB2 - 1 && pass(bk2(a + 1))) ++a;
      while (a > jc && !pass(bk2(a))) --a;
      return a;
    };
    const int jlo = first_in(inside, 400.0), jhi = last_in(inside, 400.0);
    // sum_{j in [jlo, jhi]} W2_j (-J + s (400 - b1^2 - b2_j^2 - 2 b1 b2_j c))
    const double m0 = spre[wv][0][jhi + 1] - spre[wv][0][jlo];
    const double m1 = spre[wv][1][jhi + 1] - spre[wv][1][jlo];
    const double m2 = spre[wv][2][jhi + 1] - spre[wv][2][jlo];
    const double sum = (-J + s * (400.0 - b1 * b1)) * m0 - s * (m2 + 2.0 * b1 * c * m1);
    const double w = sw1[wv][i] * kGLw5[k] * sum;
    if (POL) {
      acc_s += w * cs * cs;
      acc_p += w * (1.0 - cs * cs);
    } else {
      acc_s += w;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc_s += __shfl_xor(acc_s, o);
    if (POL) acc_p += __shfl_xor(acc_p, o);
  }
  if (lane == 0) {
    out[p] = acc_s;
    if (POL) out[npts + p] = acc_p;
  }
}

// ---------------------------------------------------------------------------
// Banded GEMM Y = What2 D on MFMA, fused with the dot Y . What1.
//
// A wave owns 16 points.  16x16x32 bf16 MFMA (or 16x16x4 f32) operand
// layout: lane l supplies row/column (l & 15) and the 8 consecutive k values
// 8 (l >> 4) .. 8 (l >> 4) + 7 of a 32-wide k step (A and B use the same k
// assignment, so the sum over k is the full product); the accumulator holds
// C[4 (l >> 4) + i][l & 15], i = 0..3.
// ---------------------------------------------------------------------------
template <int PREC, bool POL, int TG>
__global__ __launch_bounds__(64 * kMfmaWaves) void k_mfma_lumi(
    const double* __restrict__ pts, int npts, int mirror_out, FluxParams fp, MfmaGrid g,
    const uint16_t* __restrict__ Dh,
    const uint16_t* __restrict__ Dl, const float* __restrict__ D32, const uint16_t* __restrict__ Wh,
    const uint16_t* __restrict__ Wl, const float* __restrict__ W32, const double* __restrict__ wscale,
    const double* __restrict__ wsum, const int* __restrict__ ulast, int mode, const DevScalars* __restrict__ sc,
    const double* __restrict__ edge_sum, double* __restrict__ L, double* __restrict__ Ls, double* __restrict__ Lp,
    unsigned long long* __restrict__ ksteps, unsigned char* __restrict__ need, double cancel_limit)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nwork = lumi_work_items(npts, fp.mirror_ny, mirror_out);
  const int p0 = (blockIdx.x * kMfmaWaves + wv) * 16;  // first work item of the wave
  if (p0 >= nwork) return;
  auto point = [&](int q) {
    int p, pm;
    work_point(min(q, nwork - 1), fp.mirror_ny, mirror_out, &p, &pm);
    return p;
  };
  const int m = lane & 15, g4 = lane >> 4;
  double s0s = 0, s0p = 0;  // the phi sum beyond 20 fm, as in the FP64 form
  {
    const double br = mode != 1 ? sc->p20 : 1.;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if (POL) {
        const double c = cos(kPi * kGLx5[k]), sn = sin(kPi * kGLx5[k]);
        s0s += br * 1. * kGLw5[k] * c * c;
        s0p += br * 1. * kGLw5[k] * (sn * sn);
      } else {
        s0s += br * 1. * kGLw5[k];
      }
    }
  }
  const int pa = point(p0 + m);
  const long rowA = (long)k2_row(pa, npts, fp.mirror_ny) * g.nu_pad;
  // extents of the 16 rows: A (k2 rows) bounds k, the k1 rows bound u
  int kend = ulast[k2_row(pa, npts, fp.mirror_ny)] + 1;
  int uend = ulast[pa] + 1;
  for (int o = 8; o > 0; o >>= 1) {
    kend = max(kend, __shfl_xor(kend, o));
    uend = max(uend, __shfl_xor(uend, o));
  }
  kend = max(kend, __shfl_xor(kend, 16));
  kend = max(kend, __shfl_xor(kend, 32));
  uend = max(uend, __shfl_xor(uend, 16));
  uend = max(uend, __shfl_xor(uend, 32));
  const long nn = (long)g.nu_pad * g.nu_pad;
  double dot_s[4] = {0, 0, 0, 0}, dot_p[4] = {0, 0, 0, 0};
  long rowC[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) rowC[i] = (long)point(p0 + 4 * g4 + i) * g.nu_pad;
  unsigned long long nk = 0;
  const int ntiles = (uend + 15) >> 4;
  const int kbmax = (kend + 31) >> 5;
  // Column tile t (columns 16 t .. 16 t + 15) needs the k blocks
  // [kb0(t), kb1(t)) of the band.  kTileGroup consecutive tiles share one
  // load of each A fragment (the wave's 16 W2 rows at that k block), which
  // neighbouring tiles would otherwise each re-read; every tile still
  // accumulates its own k blocks in ascending order, so the grouping (the
  // tile shape the batch autotuner picks: 1, 2 or 4 tiles) does not change
  // any result.  Polarised waves carry two accumulators per tile, so they
  // group half as many tiles.
  constexpr int kTileGroup = POL ? (TG > 1 ? TG / 2 : 1) : TG;
  auto kb0 = [&](int t) { return max(0, (t * 16 - g.band + 1) >> 5); };
  auto kb1 = [&](int t) { return min(kbmax, (t * 16 + 15 + g.band - 1) / 32 + 1); };
  for (int t0 = 0; t0 < ntiles; t0 += kTileGroup) {
    f32x4 acc[kTileGroup], accp[kTileGroup];
#pragma unroll
    for (int gi = 0; gi < kTileGroup; ++gi) {
      acc[gi] = f32x4{0, 0, 0, 0};
      accp[gi] = f32x4{0, 0, 0, 0};
    }
    const int tl = min(ntiles, t0 + kTileGroup) - 1;
    for (int kb = kb0(t0); kb < kb1(tl); ++kb) {
      const int k = kb * 32 + g4 * 8;
      if (PREC == UPC_FORM_MFMA) {
        const bf16x8 ah = __builtin_bit_cast(bf16x8, *(const u16x8*)(Wh + rowA + k));
        const bf16x8 al = __builtin_bit_cast(bf16x8, *(const u16x8*)(Wl + rowA + k));
#pragma unroll
        for (int gi = 0; gi < kTileGroup; ++gi) {
          const int t = t0 + gi;
          if (t > tl || kb < kb0(t) || kb >= kb1(t)) continue;  // wave-uniform
          const long rowB = (long)(t * 16 + m) * g.nu_pad;
          const bf16x8 bh = __builtin_bit_cast(bf16x8, *(const u16x8*)(Dh + rowB + k));
          const bf16x8 bl = __builtin_bit_cast(bf16x8, *(const u16x8*)(Dl + rowB + k));
          acc[gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[gi], 0, 0, 0);
          acc[gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[gi], 0, 0, 0);
          acc[gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[gi], 0, 0, 0);
          if (POL) {
            const bf16x8 ph = __builtin_bit_cast(bf16x8, *(const u16x8*)(Dh + nn + rowB + k));
            const bf16x8 pl = __builtin_bit_cast(bf16x8, *(const u16x8*)(Dl + nn + rowB + k));
            accp[gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, ph, accp[gi], 0, 0, 0);
            accp[gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, pl, accp[gi], 0, 0, 0);
            accp[gi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, ph, accp[gi], 0, 0, 0);
          }
        }
      } else {
        const f32x8 a = *(const f32x8*)(W32 + rowA + k);
#pragma unroll
        for (int gi = 0; gi < kTileGroup; ++gi) {
          const int t = t0 + gi;
          if (t > tl || kb < kb0(t) || kb >= kb1(t)) continue;  // wave-uniform
          const long rowB = (long)(t * 16 + m) * g.nu_pad;
          const f32x8 b = *(const f32x8*)(D32 + rowB + k);
#pragma unroll
          for (int e = 0; e < 8; ++e)
            acc[gi] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e], b[e], acc[gi], 0, 0, 0);
          if (POL) {
            const f32x8 bp = *(const f32x8*)(D32 + nn + rowB + k);
#pragma unroll
            for (int e = 0; e < 8; ++e)
              accp[gi] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e], bp[e], accp[gi], 0, 0, 0);
          }
        }
      }
    }
#pragma unroll
    for (int gi = 0; gi < kTileGroup; ++gi) {
      const int t = t0 + gi;
      if (t > tl) continue;
      nk += kb1(t) > kb0(t) ? kb1(t) - kb0(t) : 0;
      const int u = t * 16 + m;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float w1;
        if (PREC == UPC_FORM_MFMA) w1 = from_bf16(Wh[rowC[i] + u]) + from_bf16(Wl[rowC[i] + u]);
        else w1 = W32[rowC[i] + u];
        dot_s[i] += (double)acc[gi][i] * (double)w1;
        if (POL) dot_p[i] += (double)accp[gi][i] * (double)w1;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    for (int o = 8; o > 0; o >>= 1) {
      dot_s[i] += __shfl_xor(dot_s[i], o);
      if (POL) dot_p[i] += __shfl_xor(dot_p[i], o);
    }
  }
  if (lane == 0) atomicAdd(ksteps, nk);
  if (m != 0) return;
  unsigned long long nflag = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int w = p0 + 4 * g4 + i;
    if (w >= nwork) continue;
    int q, qm;
    work_point(w, fp.mirror_ny, mirror_out, &q, &qm);
    const int r2 = k2_row(q, npts, fp.mirror_ny);
    const double pre = 2 * kPi * kPi * pts[4 * q] * pts[4 * q + 2] * pts[4 * q + 3];
    const double ww = wsum[q] * wsum[r2], scale = wscale[q] * wscale[r2];
    const double fs = s0s * ww, ts = fs + scale * dot_s[i] + edge_sum[q];
    bool flag = fabs(fs) > cancel_limit * fabs(ts);  // false for NaN
    if (POL) {
      const double fp_ = s0p * ww, tp = fp_ + scale * dot_p[i] + edge_sum[npts + q];
      flag = flag || fabs(fp_) > cancel_limit * fabs(tp);
      const double ls = pre * ts, lp = pre * tp;
      Ls[q] = ls;
      Lp[q] = lp;
      if (qm >= 0) {
        Ls[qm] = ls;
        Lp[qm] = lp;
      }
    } else {
      const double l = pre * ts;
      L[q] = l;
      if (qm >= 0) L[qm] = l;
    }
    need[w] = flag;
    nflag += flag;
  }
  if (nflag) atomicAdd(ksteps + 1, nflag);
}

__global__ void k_mfma_zero(unsigned long long* p)
{
  p[0] = 0ull;
  p[1] = 0ull;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

static MfmaGrid make_grid(double h, bool pol)
{
  MfmaGrid g{};
  g.h = h;
  g.inv_h = 1.0 / h;
  // extreme node cosine: the smallest b of a pair is at cos = -1 side
  const double cext = pol ? std::cos(kPi * 0.1488743389816312) : -std::cos(kPi * 0.9739065285171717);
  g.bstar = 20.0 / std::sqrt(1.0 - cext * cext) * (1.0 + 1e-12);
  g.nu = (int)std::floor(g.bstar / h) + 4;
  g.nu_pad = (g.nu + 31) / 32 * 32;
  g.band = (int)std::ceil(20.0 / h) + 1;
  g.npol = pol ? 2 : 1;
  return g;
}

static void free_buffers(MfmaState* st, bool rows_only)
{
  void* rows[] = {st->d_Wh, st->d_Wl, st->d_W32, st->d_wscale, st->d_wsum, st->d_ulast, st->d_edge_sum, st->d_need};
  for (void* p : rows)
    if (p) (void)hipFree(p);
  st->d_Wh = st->d_Wl = nullptr;
  st->d_W32 = nullptr;
  st->d_wscale = st->d_wsum = nullptr;
  st->d_ulast = nullptr;
  st->d_edge_sum = nullptr;
  st->d_need = nullptr;
  st->rows_cap = 0;
  if (rows_only) return;
  void* dm[] = {st->d_Dh, st->d_Dl, st->d_D32, st->d_ksteps, st->d_edge};
  for (void* p : dm)
    if (p) (void)hipFree(p);
  st->d_Dh = st->d_Dl = nullptr;
  st->d_D32 = nullptr;
  st->d_ksteps = nullptr;
  st->d_edge = nullptr;
  st->d_cap = 0;
}

static int ensure_buffers(MfmaState* st, size_t nrows)
{
  const MfmaGrid& g = st->g;
  const size_t dn = (size_t)g.npol * g.nu_pad * g.nu_pad;
  if (dn > st->d_cap) {
    free_buffers(st, false);
    UPC_HIP(hipMalloc(&st->d_Dh, dn * 2));
    UPC_HIP(hipMalloc(&st->d_Dl, dn * 2));
    UPC_HIP(hipMalloc(&st->d_D32, dn * 4));
    UPC_HIP(hipMalloc(&st->d_ksteps, 2 * sizeof(unsigned long long)));
    UPC_HIP(hipMalloc(&st->d_edge, 8 * sizeof(double)));
    st->d_cap = dn;
  }
  const size_t wn = nrows * g.nu_pad;
  if (wn > st->rows_cap) {
    free_buffers(st, true);
    if (st->prec == UPC_FORM_MFMA) {
      UPC_HIP(hipMalloc(&st->d_Wh, wn * 2));
      UPC_HIP(hipMalloc(&st->d_Wl, wn * 2));
    } else {
      UPC_HIP(hipMalloc(&st->d_W32, wn * 4));
    }
    UPC_HIP(hipMalloc(&st->d_wscale, nrows * sizeof(double)));
    UPC_HIP(hipMalloc(&st->d_wsum, nrows * sizeof(double)));
    UPC_HIP(hipMalloc(&st->d_ulast, nrows * sizeof(int)));
    UPC_HIP(hipMalloc(&st->d_edge_sum, 2 * nrows * sizeof(double)));
    UPC_HIP(hipMalloc(&st->d_need, nrows));
    st->rows_cap = wn;
  }
  return 0;
}

int mfma_set_spacing(upc_ctx* c, double h)
{
  if (!(h >= 0.05 && h <= 2.0)) {
    set_error("MFMA grid spacing must be in [0.05, 2] fm");
    return UPC_EINVAL;
  }
  MfmaState* st = state(c);
  st->h = h;
  return 0;
}

unsigned long long mfma_last_ksteps(upc_ctx* c) { return c->mfma ? state(c)->last[0] : 0; }
unsigned long long mfma_last_fp64_items(upc_ctx* c) { return c->mfma ? state(c)->last[1] : 0; }
int mfma_set_cancel_limit(upc_ctx* c, double limit)
{
  if (!(limit == limit)) return set_error("cancellation limit must not be NaN"), UPC_EINVAL;
  state(c)->cancel_limit = limit;
  return 0;
}
double mfma_spacing(upc_ctx* c) { return state(c)->h; }

// Buffers for views of up to nrows flux rows in precision prec; the grid
// follows the state's spacing and polarisation.
static int mfma_ready(MfmaState* st, int prec, bool pol, size_t nrows)
{
  const MfmaGrid g = make_grid(st->h, pol);
  if (prec != st->prec || g.nu_pad != st->g.nu_pad || g.npol != st->g.npol) free_buffers(st, false);
  st->prec = prec;
  st->g = g;
  return ensure_buffers(st, nrows);
}

// The MFMA contraction of one view whose flux rows are built: D from the
// view's tables (c: the job's context), the projection of its rows, the
// exact edge sums and the banded GEMM into the view's outputs, with
// tile_group column tiles per A-fragment load.  The steps count into
// st->d_ksteps (cleared by the caller).
static int mfma_contract(upc_ctx* c, MfmaState* st, const LumiView& v, int tile_group, hipStream_t s)
{
  const int prec = st->prec;
  const MfmaGrid& g = st->g;
  const bool pol = v.pol != 0;
  const int npts = v.npts;
  const FluxParams& fp = v.fp;
  const int nrows = fp.mirror_ny > 0 ? npts + npts / fp.mirror_ny : 2 * npts;
  if ((size_t)nrows * g.nu_pad > st->rows_cap) return set_error("MFMA buffers smaller than the view"), UPC_EINVAL;
  const bool split = prec == UPC_FORM_MFMA;
  const int mirror_out = v.mirror_out;
  const int nwork = lumi_work_items(npts, fp.mirror_ny, mirror_out);
  {
    PhaseScope ph(c, s, UPC_PHASE_PROJECT);
    const long dn = (long)g.nu_pad * g.nu_pad;
    uint16_t* Dh = split ? st->d_Dh : nullptr;
    uint16_t* Dl = split ? st->d_Dl : nullptr;
    float* D32 = split ? nullptr : st->d_D32;
    hipLaunchKernelGGL(k_mfma_edge_params, dim3(1), dim3(1), 0, s, v.mode, v.gaa, v.bk, v.sc, st->d_edge);
    if (pol)
      hipLaunchKernelGGL(k_mfma_kernel_matrix<true>, dim3(nblk(dn, 256)), dim3(256), 0, s, g, v.mode, v.gaa, v.bk,
                         v.sc, st->d_edge, Dh, Dl, D32);
    else
      hipLaunchKernelGGL(k_mfma_kernel_matrix<false>, dim3(nblk(dn, 256)), dim3(256), 0, s, g, v.mode, v.gaa,
                         v.bk, v.sc, st->d_edge, Dh, Dl, D32);
    hipLaunchKernelGGL(k_mfma_project, dim3(nblk(nrows, 4)), dim3(256), 0, s, nrows, v.row_logd, v.flux, fp, g,
                       split ? st->d_Wh : nullptr, split ? st->d_Wl : nullptr, split ? nullptr : st->d_W32,
                       st->d_wscale, st->d_wsum, st->d_ulast);
    if (pol)
      hipLaunchKernelGGL(k_mfma_edge<true>, dim3(nblk(nwork, 4)), dim3(256), 0, s, npts, mirror_out, v.row_logd,
                         v.flux, fp, st->d_edge, st->d_edge_sum);
    else
      hipLaunchKernelGGL(k_mfma_edge<false>, dim3(nblk(nwork, 4)), dim3(256), 0, s, npts, mirror_out, v.row_logd,
                         v.flux, fp, st->d_edge, st->d_edge_sum);
  }
  {
    PhaseScope ph(c, s, UPC_PHASE_LUMI);
    const dim3 grid(nblk(nwork, 16 * kMfmaWaves)), block(64 * kMfmaWaves);
    // limit <= 0 turns the guard off (every flag false)
    const double cancel = st->cancel_limit > 0 ? st->cancel_limit : __builtin_inf();
#define UPC_MFMA_LAUNCH(P, POLV, TGV)                                                                         \
  hipLaunchKernelGGL((k_mfma_lumi<P, POLV, TGV>), grid, block, 0, s, v.pts, npts, mirror_out, fp, g, st->d_Dh, \
                     st->d_Dl, st->d_D32, st->d_Wh, st->d_Wl, st->d_W32, st->d_wscale, st->d_wsum,             \
                     st->d_ulast, v.mode, v.sc, st->d_edge_sum, v.L, v.Ls, v.Lp, st->d_ksteps, st->d_need,   \
                     cancel)
#define UPC_MFMA_TG(P, POLV)                     \
  do {                                           \
    if (tile_group == 1) UPC_MFMA_LAUNCH(P, POLV, 1); \
    else if (tile_group == 2) UPC_MFMA_LAUNCH(P, POLV, 2); \
    else UPC_MFMA_LAUNCH(P, POLV, 4);            \
  } while (0)
    if (split) {
      if (pol) UPC_MFMA_TG(UPC_FORM_MFMA, true);
      else UPC_MFMA_TG(UPC_FORM_MFMA, false);
    } else {
      if (pol) UPC_MFMA_TG(UPC_FORM_MFMA_FP32, true);
      else UPC_MFMA_TG(UPC_FORM_MFMA_FP32, false);
    }
#undef UPC_MFMA_TG
#undef UPC_MFMA_LAUNCH
    UPC_HIP(hipGetLastError());
    // the work items the cancellation guard flagged, in FP64 from the same flux rows
    if (st->cancel_limit > 0) {
      LumiView vf = v;
      vf.need = st->d_need;
      if (int rc = launch_lumi_fp64(vf, s)) return rc;
    }
  }
  return 0;
}

int lumi_points_mfma(upc_ctx* c, int prec, const double* d_pts, int npts, double* d_L, double* d_Ls,
                     double* d_Lp, hipStream_t s)
{
  if (npts <= 0) return 0;
  MfmaState* st = state(c);
  FluxParams fp;
  int rc = prepare_flux(c, d_pts, npts, &fp, s);
  if (rc) return rc;
  const int nrows = fp.mirror_ny > 0 ? npts + npts / fp.mirror_ny : 2 * npts;
  rc = mfma_ready(st, prec, c->cfg.use_polarized != 0, (size_t)nrows);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mfma_zero, dim3(1), dim3(1), 0, s, st->d_ksteps);
  LumiView v{c->work.d_points, c->work.d_row_logd, c->work.d_flux, fp, c->d_gaa, c->d_bk, c->d_scalars, npts,
             c->cfg.breakup_mode, c->cfg.use_polarized, d_L, d_Ls, d_Lp, fp.mirror_ny > 0};
  rc = mfma_contract(c, st, v, kMfmaTileGroup, s);
  if (rc) return rc;
  UPC_HIP(hipMemcpyAsync(st->last, st->d_ksteps, sizeof(st->last), hipMemcpyDeviceToHost, s));
  return read_flux_stats(c, s);
}

// Batch use: a state of its own (spacing h), sized for views of up to
// nrows flux rows.
void* mfma_state_create(double h) { auto* st = new MfmaState(); st->h = h; return st; }
void mfma_state_free(void* p)
{
  if (!p) return;
  auto* st = static_cast<MfmaState*>(p);
  free_buffers(st, false);
  delete st;
}
int mfma_state_ready(void* p, int prec, bool pol, size_t nrows)
{
  return mfma_ready(static_cast<MfmaState*>(p), prec, pol, nrows);
}
int mfma_batch_contract(upc_ctx* c, void* p, const LumiView& v, int tile_group, hipStream_t s)
{
  return mfma_contract(c, static_cast<MfmaState*>(p), v, tile_group, s);
}
void mfma_zero_ksteps(void* p, hipStream_t s)
{
  hipLaunchKernelGGL(k_mfma_zero, dim3(1), dim3(1), 0, s, static_cast<MfmaState*>(p)->d_ksteps);
}
// MFMA tile steps and work items recomputed in FP64 since the state's last
// mfma_zero_ksteps (copied after the work queued on s).
int mfma_read_counts(void* p, unsigned long long* h_out, hipStream_t s)
{
  auto* st = static_cast<MfmaState*>(p);
  h_out[0] = h_out[1] = 0;
  if (!st->d_ksteps) return 0;
  UPC_HIP(hipMemcpyAsync(h_out, st->d_ksteps, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipStreamSynchronize(s));
  return 0;
}

int lumi_grid_mfma(upc_ctx* c, int prec, double* d_L, double* d_Ls, double* d_Lp, hipStream_t s)
{
  int rc = make_grid_points(c, s);
  if (rc) return rc;
  return lumi_points_mfma(c, prec, nullptr, c->cfg.nm * c->cfg.ny, d_L, d_Ls, d_Lp, s);
}

void free_mfma(upc_ctx* c)
{
  if (!c->mfma) return;
  MfmaState* st = state(c);
  free_buffers(st, false);
  delete st;
  c->mfma = nullptr;
}

}  // namespace upc
