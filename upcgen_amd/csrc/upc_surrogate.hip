// MLP surrogate of the two-photon luminosity, trained and served by the
// kernels in this file.
//
// No reference counterpart: the reference evaluates every L(M, Y) by
// quadrature (calcTwoPhotonLumi, src/UpcCrossSection.cpp:221-271).  The
// surrogate predicts
//   r = log L(M, Y; system) - log L0,   L0 = (M / 2) n(k1) n(k2),
// where n(k) = 2 pi int_R^inf b N_point(b, k) db
//           = (2 Z^2 alpha / (pi k)) [xi K0 K1 - xi^2/2 (K1^2 - K0^2)],  xi = k R / (gamma hc)
// is the point-charge photon number beyond b = R (the reference's L with P =
// G = 1 is (M / 2) n(k1) n(k2) up to its b cut-offs).
//
// Network: inputs (log M, Y, log gamma, Z, A^(1/3), R, a, one-hot breakup
// mode, log xi1, log xi2; padded to 32), 3 hidden layers of 128 SiLU units,
// one output.
// Forward and backward products run on MFMA (v_mfma_f32_16x16x32_bf16) with
// bf16 operands rounded from the FP32 master copy and FP32 accumulation; the
// output layer is FP32 with bf16-rounded weights.  Adam on FP32 master
// weights.  Two networks with different seeds form an ensemble: their mean
// is the prediction and their difference the predicted error, and points
// whose predicted error exceeds the caller's bound are recomputed by the FP64
// quadrature.
#include "upc_flux.h"
#include "upc_philox.h"

#include <cmath>
#include <cstring>
#include <vector>

struct upc_sur;

namespace upc {

constexpr int kSurIn = 32;     // padded input width
constexpr int kSurFeat = 13;   // features in use
constexpr int kSurH = 128;
constexpr int kSurNets = 2;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kSurH * kSurIn;
constexpr int kOffW2 = kOffB1 + kSurH;
constexpr int kOffB2 = kOffW2 + kSurH * kSurH;
constexpr int kOffW3 = kOffB2 + kSurH;
constexpr int kOffB3 = kOffW3 + kSurH * kSurH;
constexpr int kOffW4 = kOffB3 + kSurH;
constexpr int kOffB4 = kOffW4 + kSurH;
constexpr int kSurParams = kOffB4 + 1;
constexpr int kSurMaxBatch = 65536;  // rows of the activation buffers
// Domain of the surrogate: both photons with xi = k R / (gamma hc) <= 4.
// Beyond it L falls like exp(-2 xi) and the reference's QAGS values
// (epsabs 1e-4) are mostly quadrature error; those bins always go to the
// quadrature (log L0 is set to NaN there).
constexpr double kSurXiMax = 4.0;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short bf16_bits(float f)
{
  const uint32_t u = __float_as_uint(f);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite input)
}
__device__ __forceinline__ float bf16_value(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float bf16_round(float f) { return bf16_value(bf16_bits(f)); }

enum { kEpiStore = 0, kEpiBiasSilu = 1 };

// C[M x N] = opA(A)[M x K] opB(B)[K x N], bf16 operands, FP32 accumulation.
// TA: element (m, k) of A at A[k * lda + m], else A[m * lda + k].
// TB: element (k, n) of B at B[n * ldb + k], else B[k * ldb + n].
// M, N multiples of 16, K a multiple of 32.  One wave per 16 x 16 tile with
// the 16x16x32 MFMA layout: lane l supplies row / column (l & 15) and k
// values 8 (l >> 4) .. +7 of each 32-deep step, and holds C[4 (l >> 4) + i][l & 15].
// kEpiBiasSilu: C = acc + bias[n] (pre-activation) and H = silu(C).
template <bool TA, bool TB, int EPI, bool SPLIT>
__global__ __launch_bounds__(256) void k_sur_mm(int M, int N, int K, const float* __restrict__ A, int lda,
                                                const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                int ldc, const float* __restrict__ bias, float* __restrict__ H)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles_n = N >> 4;
  const int tile = blockIdx.x * 4 + wv;
  if (tile >= (M >> 4) * tiles_n) return;
  const int m0 = (tile / tiles_n) << 4, n0 = (tile % tiles_n) << 4;
  const int r = lane & 15, g4 = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    u16x8 a, b, al, bl;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + g4 * 8 + e;
      const float av = TA ? A[(long)k * lda + m0 + r] : A[(long)(m0 + r) * lda + k];
      const float bv = TB ? B[(long)(n0 + r) * ldb + k] : B[(long)k * ldb + n0 + r];
      a[e] = bf16_bits(av);
      b[e] = bf16_bits(bv);
      if (SPLIT) {  // the remainders (exact in FP32) rounded to bf16 again
        al[e] = bf16_bits(av - bf16_value(a[e]));
        bl[e] = bf16_bits(bv - bf16_value(b[e]));
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                  acc, 0, 0, 0);
    if (SPLIT) {
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, bl),
                                                    acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, al), __builtin_bit_cast(bf16x8, b),
                                                    acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long m = m0 + 4 * g4 + i, n = n0 + r;
    float v = acc[i];
    if (EPI == kEpiBiasSilu) {
      v += bias[n];
      C[m * ldc + n] = v;
      H[m * ldc + n] = v / (1.f + expf(-v));
    } else {
      C[m * ldc + n] = v;
    }
  }
}

// Output layer: pred[b] = sum_o w4(W4[o]) H3[b][o] + b4, w4 = bf16 rounding
// (plain) or the FP32 weight (split); with targets, the loss gradient
// g[b] = 2 (pred - t) / B and the summed squared error.
template <bool SPLIT>
__device__ __forceinline__ float out_weight(float w) { return SPLIT ? w : bf16_round(w); }

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_sur_out(int B, const float* __restrict__ H3, const float* __restrict__ p,
                                                 const float* __restrict__ t, float* __restrict__ pred,
                                                 float* __restrict__ g, float* __restrict__ loss,
                                                 const int* __restrict__ step)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + wv;
  if (b >= B) return;
  float s = out_weight<SPLIT>(p[kOffW4 + lane]) * H3[(long)b * kSurH + lane] +
            out_weight<SPLIT>(p[kOffW4 + 64 + lane]) * H3[(long)b * kSurH + 64 + lane];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  s += p[kOffB4];
  if (lane == 0) {
    pred[b] = s;
    if (t) {
      const float d = s - t[b];
      g[b] = 2.f * d / B;
      if (loss) atomicAdd(&loss[step ? *step : 0], d * d / B);
    }
  }
}

// Inference of one network in one launch (no activations kept for a
// backward pass): the three hidden layers and the output for 16 rows at a
// time per wavefront, with the weights in LDS and the activations in
// registers (212 VGPRs split, 169 plain; two row tiles at once would spill).
//
// The layers are computed transposed, C'[n][b] = sum_k W[n][k] H[b][k], so
// the weights are the MFMA's A operand (16 output units per tile) and the
// activations its B operand (16 rows per tile).  The 16x16x32 layout leaves
// lane (r, g) = (l & 15, l >> 4) holding units 16 t + 4 g + i (i < 4) of row
// r for each unit tile t; as the B operand of the next layer's k-step s the
// lane supplies exactly those it holds, units 32 s + 4 g + j and
// 32 s + 16 + 4 g + j (j < 4), so the weights are stored with their k index
// in the same order and no activation crosses a lane between layers.
// Products are split bf16 (w_hi h_hi + w_hi h_lo + w_lo h_hi) or plain bf16
// as in k_sur_mm, with FP32 accumulation and SiLU; the output layer is FP32
// with bf16-rounded weights (plain) or FP32 weights (split).  Only the order
// of the FP32 sums differs from the layer-by-layer kernels.
//
// LDS: per layer, unit tile t and k-step s, 64 lanes x 8 bf16 (hi, then lo),
// 147,456 B for the three layers, so one workgroup of 8 wavefronts per CU
// (2 per SIMD).  k_sur_infer_prep writes that image once per call and every
// workgroup copies it in with 16-byte loads.
constexpr int kInferWaves = 8;
constexpr int kInferTiles = 1;  // row tiles of 16 per wavefront

__device__ __forceinline__ int infer_k(int s, int g, int j) { return 32 * s + 4 * g + (j < 4 ? j : 12 + j); }

constexpr int kInferL1 = (kSurH / 16) * 64 * 8;             // layer-1 A operands (one k-step)
constexpr int kInferL23 = (kSurH / 16) * (kSurH / 32) * 64 * 8;  // layers 2, 3
constexpr int kInferWords = 2 * kInferL1 + 4 * kInferL23;       // bf16 hi and lo of the three layers

// The A-operand image of one network's weights: per layer, unit tile t and
// k-step s, lane (r, g) of the 16x16x32 layout holds W[16 t + r][k] for its 8
// k (layer 1: k = 8 g + j; layers 2, 3: k = infer_k(s, g, j)), as bf16 hi
// and lo: w1 hi, w1 lo, w2 hi, w2 lo, w3 hi, w3 lo.
__global__ void k_sur_infer_prep(const float* __restrict__ p, unsigned short* __restrict__ img)
{
  constexpr int kS = kSurH / 32, kBlk = 64 * 8;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kInferL1 + 2 * kInferL23) return;
  const int layer = e < kInferL1 ? 0 : 1 + (e - kInferL1) / kInferL23;
  const int le = layer == 0 ? e : (e - kInferL1) % kInferL23;
  const int blk = le / kBlk, lane = (le % kBlk) >> 3, j = le & 7;
  const int t = layer == 0 ? blk : blk / kS, s = layer == 0 ? 0 : blk % kS;
  const int n = 16 * t + (lane & 15), g = lane >> 4;
  const float w = layer == 0 ? p[kOffW1 + n * kSurIn + 8 * g + j]
                             : p[(layer == 1 ? kOffW2 : kOffW3) + n * kSurH + infer_k(s, g, j)];
  const unsigned short hi = bf16_bits(w);
  const unsigned short lo = bf16_bits(w - bf16_value(hi));
  unsigned short* base = layer == 0 ? img : img + 2 * kInferL1 + 2 * (layer - 1) * kInferL23;
  const int half = layer == 0 ? kInferL1 : kInferL23;
  base[le] = hi;
  base[half + le] = lo;
}

template <bool SPLIT>
__global__ __launch_bounds__(64 * kInferWaves) void k_sur_infer(int B, const float* __restrict__ X,
                                                                 const float* __restrict__ p,
                                                                 const unsigned short* __restrict__ img,
                                                                 float* __restrict__ pred)
{
  constexpr int kT = kSurH / 16;           // unit tiles
  constexpr int kS = kSurH / 32;           // k-steps of layers 2 and 3
  constexpr int kBlk = 64 * 8;             // bf16 per (tile, step) block
  __shared__ __align__(16) unsigned short wimg[kInferWords];
  __shared__ float bias[3][kSurH], w4[kSurH];
  const unsigned short* w1h = wimg;
  const unsigned short* w1l = wimg + kInferL1;
  // the prepared image, 16 bytes per load, all of a thread's loads in flight
  {
    constexpr int kChunks = kInferWords / 8;
    static_assert(kChunks % (64 * kInferWaves) == 0, "whole chunks per thread");
    constexpr int kPer = kChunks / (64 * kInferWaves);
    uint4 v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) v[k] = ((const uint4*)img)[k * 64 * kInferWaves + threadIdx.x];
#pragma unroll
    for (int k = 0; k < kPer; ++k) ((uint4*)wimg)[k * 64 * kInferWaves + threadIdx.x] = v[k];
  }
  for (int e = threadIdx.x; e < kSurH; e += blockDim.x) {
    bias[0][e] = p[kOffB1 + e];
    bias[1][e] = p[kOffB2 + e];
    bias[2][e] = p[kOffB3 + e];
    w4[e] = out_weight<SPLIT>(p[kOffW4 + e]);
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const float b4 = p[kOffB4];
  const int ntiles = (B + 15) >> 4;
  for (int t0 = (blockIdx.x * kInferWaves + wv) * kInferTiles; t0 < ntiles;
       t0 += gridDim.x * kInferWaves * kInferTiles) {
    f32x4 h[kInferTiles][kT];
    // layer 1: B operand = the rows' 32 features, k = 8 g + j
    {
      bf16x8 xh[kInferTiles], xl[kInferTiles];
#pragma unroll
      for (int q = 0; q < kInferTiles; ++q) {
        const int row = (t0 + q) * 16 + r;
        u16x8 a, al;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const float x = row < B ? X[(long)row * kSurIn + 8 * g + j] : 0.f;
          a[j] = bf16_bits(x);
          al[j] = bf16_bits(x - bf16_value(a[j]));
        }
        xh[q] = __builtin_bit_cast(bf16x8, a);
        xl[q] = __builtin_bit_cast(bf16x8, al);
      }
#pragma unroll
      for (int t = 0; t < kT; ++t) {
        const bf16x8 wa = __builtin_bit_cast(bf16x8, *(const u16x8*)&w1h[t * kBlk + lane * 8]);
        const bf16x8 wl = __builtin_bit_cast(bf16x8, *(const u16x8*)&w1l[t * kBlk + lane * 8]);
#pragma unroll
        for (int q = 0; q < kInferTiles; ++q) {
          f32x4 acc = {0.f, 0.f, 0.f, 0.f};
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xh[q], acc, 0, 0, 0);
          if (SPLIT) {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xl[q], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl, xh[q], acc, 0, 0, 0);
          }
          h[q][t] = acc;
        }
      }
#pragma unroll
      for (int q = 0; q < kInferTiles; ++q)
#pragma unroll
        for (int t = 0; t < kT; ++t)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float v = h[q][t][i] + bias[0][16 * t + 4 * g + i];
            h[q][t][i] = v / (1.f + expf(-v));
          }
    }
    // layers 2 and 3
#pragma unroll 1
    for (int layer = 0; layer < 2; ++layer) {
      f32x4 o[kInferTiles][kT];
#pragma unroll
      for (int q = 0; q < kInferTiles; ++q)
#pragma unroll
        for (int t = 0; t < kT; ++t) o[q][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < kS; ++s) {
        bf16x8 bh[kInferTiles], bl[kInferTiles];
#pragma unroll
        for (int q = 0; q < kInferTiles; ++q) {
          u16x8 a, al;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const float x = h[q][2 * s + (j >> 2)][j & 3];
            a[j] = bf16_bits(x);
            al[j] = bf16_bits(x - bf16_value(a[j]));
          }
          bh[q] = __builtin_bit_cast(bf16x8, a);
          bl[q] = __builtin_bit_cast(bf16x8, al);
        }
#pragma unroll
        for (int t = 0; t < kT; ++t) {
          const int off = 2 * kInferL1 + 2 * layer * kInferL23 + (t * kS + s) * kBlk + lane * 8;
          const bf16x8 wa = __builtin_bit_cast(bf16x8, *(const u16x8*)&wimg[off]);
          const bf16x8 wl = __builtin_bit_cast(bf16x8, *(const u16x8*)&wimg[off + kInferL23]);
#pragma unroll
          for (int q = 0; q < kInferTiles; ++q) {
            o[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, bh[q], o[q][t], 0, 0, 0);
            if (SPLIT) {
              o[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, bl[q], o[q][t], 0, 0, 0);
              o[q][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl, bh[q], o[q][t], 0, 0, 0);
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kInferTiles; ++q)
#pragma unroll
        for (int t = 0; t < kT; ++t)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float v = o[q][t][i] + bias[1 + layer][16 * t + 4 * g + i];
            h[q][t][i] = v / (1.f + expf(-v));
          }
    }
    // output: the lane's 32 units, then the 4 lanes of each row
#pragma unroll
    for (int q = 0; q < kInferTiles; ++q) {
      float sum = 0.f;
#pragma unroll
      for (int t = 0; t < kT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) sum += w4[16 * t + 4 * g + i] * h[q][t][i];
      sum += __shfl_xor(sum, 16);
      sum += __shfl_xor(sum, 32);
      const int row = (t0 + q) * 16 + r;
      if (g == 0 && row < B) pred[row] = sum + b4;
    }
  }
}

// dZ3 = g[b] w4(W4[o]) silu'(Z3)
template <bool SPLIT>
__global__ void k_sur_out_bwd(int B, const float* __restrict__ g, const float* __restrict__ p,
                              const float* __restrict__ Z, float* __restrict__ dZ)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * kSurH) return;
  const int o = (int)(i % kSurH);
  const float z = Z[i], s = 1.f / (1.f + expf(-z));
  dZ[i] = g[i / kSurH] * out_weight<SPLIT>(p[kOffW4 + o]) * (s * (1.f + z * (1.f - s)));
}

// dZ = dH silu'(Z), in place on dH
__global__ void k_sur_silu_bwd(long n, const float* __restrict__ Z, float* __restrict__ dH)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float z = Z[i], s = 1.f / (1.f + expf(-z));
  dH[i] *= s * (1.f + z * (1.f - s));
}

// out[n] = sum_b w[b] X[b][n] (w = 1 when null): one block per column, fixed order.
__global__ __launch_bounds__(256) void k_sur_colsum(int B, int N, const float* __restrict__ X,
                                                    const float* __restrict__ w, float* __restrict__ out)
{
  __shared__ float red[256];
  const int n = blockIdx.x;
  float s = 0.f;
  for (int b = threadIdx.x; b < B; b += 256) s += (w ? w[b] : 1.f) * X[(long)b * N + n];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = red[0];
}

__global__ void k_sur_sum(int B, const float* __restrict__ g, float* __restrict__ out)
{
  __shared__ float red[256];
  float s = 0.f;
  for (int b = threadIdx.x; b < B; b += 256) s += g[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Training-step bookkeeping on the device so a captured step replays as is.
__global__ void k_sur_step_inc(int* step) { ++*step; }
__global__ void k_batch_zero_i(int* p) { *p = 0; }
__global__ void k_batch_zero_f(float* p, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

struct AdamCfg {
  float lr, lr_min, b1, b2, eps;
  int total;
};

// Adam with a cosine learning-rate decay over `total` steps; step t = *step (1-based).
__global__ void k_sur_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, AdamCfg a, const int* __restrict__ step)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = *step;
  const float frac = fminf(1.f, (float)(t - 1) / (float)(a.total > 1 ? a.total - 1 : 1));
  const float lr = a.lr_min + 0.5f * (a.lr - a.lr_min) * (1.f + cosf(3.14159265358979f * frac));
  const float gi = g[i];
  const float mi = a.b1 * m[i] + (1.f - a.b1) * gi;
  const float vi = a.b2 * v[i] + (1.f - a.b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(a.b1, (float)t));
  const float vh = vi / (1.f - powf(a.b2, (float)t));
  p[i] -= lr * mh / (sqrtf(vh) + a.eps);
}

// Mini-batch b of step t: rows drawn with Philox (seed, t * B + b).
__global__ void k_sur_gather(int B, long n, unsigned long long seed, const int* __restrict__ step,
                             const float* __restrict__ X, const float* __restrict__ T, float* __restrict__ Xb,
                             float* __restrict__ Tb)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  Philox4x32 ph;
  ph.block(seed, (unsigned long long)(*step) * (unsigned long long)B + b, 0);
  const unsigned long long u = ((unsigned long long)ph.c[0] << 32) | ph.c[1];
  const long row = (long)(u % (unsigned long long)n);
  for (int j = 0; j < kSurIn; ++j) Xb[(long)b * kSurIn + j] = X[row * kSurIn + j];
  Tb[b] = T[row];
}

// Features and the analytic baseline log L0 of (M, Y) for one system.
struct SurSystem {
  double Z, A, R, a, gamma;
  int mode;
};

__device__ __forceinline__ double log_photon_number(double k, const SurSystem& s)
{
  const double xi = k * s.R / (s.gamma * kHc);
  double k0s, k1s;  // e^xi K0, e^xi K1
  bessel_k01_scaled(xi, &k0s, &k1s);
  const double br = xi * k0s * k1s - 0.5 * xi * xi * (k1s * k1s - k0s * k0s);
  return log(2.0 * s.Z * s.Z * kAlpha / (kPi * k)) - 2.0 * xi + log(br);
}

__global__ void k_sur_features(SurSystem s, const double* __restrict__ Mv, const double* __restrict__ Yv,
                               long n, int stride, float* __restrict__ X, double* __restrict__ logL0)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double M = Mv[i * stride], Y = Yv[i * stride];
  float* x = X + i * kSurIn;
  x[0] = (float)((log(M) - 2.65) / 1.53);
  x[1] = (float)(Y / 4.0);
  x[2] = (float)((log(s.gamma) - 7.3) / 1.55);
  x[3] = (float)((s.Z - 50.0) / 24.5);
  x[4] = (float)(cbrt(s.A) - 4.8);
  x[5] = (float)((s.R - 5.5) / 1.5);
  x[6] = (float)((s.a - 0.5) / 0.05);
  for (int m = 0; m < 4; ++m) x[7 + m] = s.mode == m + 1 ? 1.f : 0.f;
  const double k1 = M / 2. * exp(Y), k2 = M / 2. * exp(-Y);
  const double xi1 = k1 * s.R / (s.gamma * kHc), xi2 = k2 * s.R / (s.gamma * kHc);
  x[11] = (float)((log(xi1) + 3.0) / 3.0);
  x[12] = (float)((log(xi2) + 3.0) / 3.0);
  for (int j = kSurFeat; j < kSurIn; ++j) x[j] = 0.f;
  const double xmax = fmax(xi1, xi2);
  logL0[i] = xmax <= kSurXiMax ? log(M / 2.) + log_photon_number(k1, s) + log_photon_number(k2, s) : NAN;
}

SurSystem sur_system(const upc_config& c)
{
  SurSystem s;
  s.Z = c.Z;
  s.A = c.A;
  s.R = c.R;
  s.a = c.a;
  s.gamma = c.sqrts / (2. * kMProt);
  s.mode = c.breakup_mode;
  return s;
}

// Grid form: accept the ensemble where its spread is within the bound.
__global__ void k_sur_combine(long n, const float* __restrict__ pa, const float* __restrict__ pb,
                              const double* __restrict__ logL0, double mu, double sigma, double err_scale,
                              double max_err,
                              const double* __restrict__ pts, double* __restrict__ L, int* __restrict__ fallback,
                              unsigned long long* __restrict__ nfall, double* __restrict__ err_out)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ra = mu + sigma * pa[i], rb = mu + sigma * pb[i];
  const double err = err_scale * fabs(ra - rb);
  const double lg = logL0[i] + 0.5 * (ra + rb);
  if (err_out) err_out[i] = err;
  if (isfinite(lg) && isfinite(logL0[i]) && err <= max_err) {
    L[i] = exp(lg) * pts[4 * i + 2] * pts[4 * i + 3];
  } else {
    const unsigned long long slot = atomicAdd(nfall, 1ull);
    fallback[slot] = (int)i;
  }
}

// Anchor lattice of the grid form's error check: along an axis of n bins
// with spacing `every`, the bins 0, every, 2 every, ... and n - 1.
__host__ __device__ __forceinline__ int anchor_count(int n, int every)
{
  return (n - 1) / every + 1 + ((n - 1) % every != 0 ? 1 : 0);
}
__host__ __device__ __forceinline__ int anchor_pos(int a, int n, int every)
{
  const int p = a * every;
  return p < n ? p : n - 1;
}
// Anchor index of bin i on its axis, or -1.
__device__ __forceinline__ int anchor_of(int i, int n, int every, int na)
{
  return i % every == 0 ? i / every : (i == n - 1 ? na - 1 : -1);
}

struct AnchorGrid {
  int nm, ny, em, ey, nam, nay;
};

// The network's |d log L| at each anchor against the FP64 quadrature value
// La there (infinite where the anchor lies outside the domain or either
// value is not finite).
__global__ void k_sur_anchor_err(AnchorGrid g, const float* __restrict__ pa, const float* __restrict__ pb,
                                 const double* __restrict__ logL0, double mu, double sigma,
                                 const double* __restrict__ La, double* __restrict__ aerr)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= g.nam * g.nay) return;
  const int am = a / g.nay, ay = a - am * g.nay;
  const long p = (long)anchor_pos(am, g.nm, g.em) * g.ny + anchor_pos(ay, g.ny, g.ey);
  const double lg = logL0[p] + 0.5 * ((mu + sigma * pa[p]) + (mu + sigma * pb[p]));
  const double la = La[a];
  aerr[a] = isfinite(lg) && la > 0.0 && isfinite(la) ? fabs(lg - log(la)) : INFINITY;
}

// Bins that the FP64 quadrature answers before the gate: the anchors and
// the bins outside the network's domain (log L0 not finite).  *count
// receives their number.
__global__ void k_sur_need_before(AnchorGrid g, const double* __restrict__ logL0, unsigned char* __restrict__ need,
                                  unsigned long long* __restrict__ count)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)g.nm * g.ny) return;
  const int im = (int)(i / g.ny), iy = (int)(i - (long)im * g.ny);
  const bool anchor = anchor_of(im, g.nm, g.em, g.nam) >= 0 && anchor_of(iy, g.ny, g.ey, g.nay) >= 0;
  const bool v = anchor || !isfinite(logL0[i]);
  need[i] = v;
  if (v) atomicAdd(count, 1ull);
}

// L at each anchor from the FP64 grid values (stored x dm dy).
__global__ void k_sur_anchor_vals(AnchorGrid g, const double* __restrict__ Lq, const double* __restrict__ pts,
                                  double* __restrict__ La)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= g.nam * g.nay) return;
  const int am = a / g.nay, ay = a - am * g.nay;
  const long p = (long)anchor_pos(am, g.nm, g.em) * g.ny + anchor_pos(ay, g.ny, g.ey);
  La[a] = Lq[p] / (pts[4 * p + 2] * pts[4 * p + 3]);
}

// Grid gate with anchors: a bin answered before the gate (need_before: an
// anchor, or outside the domain) takes its quadrature value Lq; any other
// bin takes the network's value when the largest anchor error at the four
// corners of its anchor cell is within margin * max_err, and falls back to
// the quadrature otherwise.
__global__ void k_sur_gate(AnchorGrid g, const float* __restrict__ pa, const float* __restrict__ pb,
                           const double* __restrict__ logL0, double mu, double sigma,
                           const double* __restrict__ aerr, const unsigned char* __restrict__ need_before,
                           const double* __restrict__ Lq, double bound, const double* __restrict__ pts,
                           double* __restrict__ L, int* __restrict__ fallback, unsigned long long* __restrict__ nfall)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)g.nm * g.ny) return;
  const int im = (int)(i / g.ny), iy = (int)(i - (long)im * g.ny);
  const double dmdy = pts[4 * i + 2] * pts[4 * i + 3];
  if (need_before[i]) {
    L[i] = Lq[i];
    return;
  }
  const int m0 = im / g.em, m1 = m0 + 1 < g.nam ? m0 + 1 : g.nam - 1;
  const int y0 = iy / g.ey, y1 = y0 + 1 < g.nay ? y0 + 1 : g.nay - 1;
  const double e = fmax(fmax(aerr[m0 * g.nay + y0], aerr[m0 * g.nay + y1]),
                        fmax(aerr[m1 * g.nay + y0], aerr[m1 * g.nay + y1]));
  const double lg = logL0[i] + 0.5 * ((mu + sigma * pa[i]) + (mu + sigma * pb[i]));
  if (isfinite(lg) && e <= bound) {
    L[i] = exp(lg) * dmdy;
  } else {
    const unsigned long long slot = atomicAdd(nfall, 1ull);
    fallback[slot] = (int)i;
  }
}

__global__ void k_sur_mark(long n, const int* __restrict__ idx, unsigned char* __restrict__ need)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  need[idx[i]] = 1;
}

__global__ void k_sur_take(long n, const int* __restrict__ idx, const double* __restrict__ Lg,
                           double* __restrict__ L)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  L[idx[i]] = Lg[idx[i]];
}

static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace upc

using namespace upc;

struct upc_sur {
  float* d_p = nullptr;   // [nets][P] master weights
  float* d_m = nullptr;   // Adam moments
  float* d_v = nullptr;
  float* d_g = nullptr;   // [P] gradient of the current step
  float* d_act = nullptr; // Z1 H1 Z2 H2 Z3 H3 [6][max_batch][128]
  float* d_xb = nullptr;  // [max_batch][32]
  unsigned short* d_img = nullptr;  // k_sur_infer's weight image of one network (prepared per call)
  float* d_tb = nullptr;
  float* d_pred = nullptr;
  float* d_gout = nullptr;
  float* d_loss = nullptr;
  int loss_cap = 0;
  int* d_step = nullptr;
  double mu = 0, sigma = 1;
  double err_scale = 1;  // predicted error = err_scale * |r_a - r_b| (calibrated on held-out data)
  int split = 1;         // products as split bf16 (hi hi + hi lo + lo hi), else plain bf16
  unsigned long long seed = 1;
  hipStream_t cap_stream = nullptr;
};

namespace {

float* act(upc_sur* s, int which) { return s->d_act + (long)which * kSurMaxBatch * kSurH; }

// C = opA(A) opB(B) (+ bias, SiLU) in the surrogate's product precision.
template <bool TA, bool TB, int EPI>
void mm(const upc_sur* s, int tiles, hipStream_t st, int M, int N, int K, const float* A, int lda, const float* B,
        int ldb, float* C, int ldc, const float* bias, float* H)
{
  if (s->split)
    hipLaunchKernelGGL((k_sur_mm<TA, TB, EPI, true>), dim3(nblk(tiles, 4)), dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, bias, H);
  else
    hipLaunchKernelGGL((k_sur_mm<TA, TB, EPI, false>), dim3(nblk(tiles, 4)), dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, bias, H);
}

// Prediction of net `net` for B rows of X (one fused launch).
int infer(upc_sur* s, int net, const float* X, int B, float* pred, hipStream_t st)
{
  if (B <= 0) return 0;
  const float* p = s->d_p + (long)net * kSurParams;
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int per_blk = kInferWaves * kInferTiles;  // row tiles of a workgroup's pass
  const int blocks = std::max(1, std::min(ncu > 0 ? ncu : 256, nblk(((long)B + 15) / 16, per_blk)));
  hipLaunchKernelGGL(k_sur_infer_prep, dim3(nblk(kInferL1 + 2 * kInferL23, 256)), dim3(256), 0, st, p, s->d_img);
  if (s->split)
    hipLaunchKernelGGL(k_sur_infer<true>, dim3(blocks), dim3(64 * kInferWaves), 0, st, B, X, p, s->d_img, pred);
  else
    hipLaunchKernelGGL(k_sur_infer<false>, dim3(blocks), dim3(64 * kInferWaves), 0, st, B, X, p, s->d_img, pred);
  UPC_HIP(hipGetLastError());
  return 0;
}

// forward of net `net` on B rows of s->d_xb (or X); activations in d_act.
// Without targets (prediction only) it is the fused inference kernel.
int forward(upc_sur* s, int net, const float* X, int B, const float* T, float* pred, float* loss, hipStream_t st)
{
  if (!T && !loss) return infer(s, net, X, B, pred, st);
  const float* p = s->d_p + (long)net * kSurParams;
  float *Z1 = act(s, 0), *H1 = act(s, 1), *Z2 = act(s, 2), *H2 = act(s, 3), *Z3 = act(s, 4), *H3 = act(s, 5);
  const int tiles = (B / 16) * (kSurH / 16);
  mm<false, true, kEpiBiasSilu>(s, tiles, st, B, kSurH, kSurIn,
                     X, kSurIn, p + kOffW1, kSurIn, Z1, kSurH, p + kOffB1, H1);
  mm<false, true, kEpiBiasSilu>(s, tiles, st, B, kSurH, kSurH,
                     H1, kSurH, p + kOffW2, kSurH, Z2, kSurH, p + kOffB2, H2);
  mm<false, true, kEpiBiasSilu>(s, tiles, st, B, kSurH, kSurH,
                     H2, kSurH, p + kOffW3, kSurH, Z3, kSurH, p + kOffB3, H3);
  if (s->split)
    hipLaunchKernelGGL(k_sur_out<true>, dim3(nblk(B, 4)), dim3(256), 0, st, B, H3, p, T, pred, s->d_gout, loss,
                       s->d_step);
  else
    hipLaunchKernelGGL(k_sur_out<false>, dim3(nblk(B, 4)), dim3(256), 0, st, B, H3, p, T, pred, s->d_gout, loss,
                       s->d_step);
  UPC_HIP(hipGetLastError());
  return 0;
}

// gradient of the mean squared error of the last forward into s->d_g
int backward(upc_sur* s, int net, const float* X, int B, hipStream_t st)
{
  const float* p = s->d_p + (long)net * kSurParams;
  float* g = s->d_g;
  float *Z1 = act(s, 0), *H1 = act(s, 1), *Z2 = act(s, 2), *H2 = act(s, 3), *Z3 = act(s, 4), *H3 = act(s, 5);
  // reuse Z buffers' partners for dZ: dZ3 -> buffer 6? keep separate scratch after the activations
  float* dA = s->d_act + 6L * kSurMaxBatch * kSurH;  // dZ / dH scratch, 2 buffers
  float* dB = dA + (long)kSurMaxBatch * kSurH;
  const long n = (long)B * kSurH;
  hipLaunchKernelGGL(k_sur_colsum, dim3(kSurH), dim3(256), 0, st, B, kSurH, H3, s->d_gout, g + kOffW4);
  hipLaunchKernelGGL(k_sur_sum, dim3(1), dim3(256), 0, st, B, s->d_gout, g + kOffB4);
  if (s->split)  // dZ3
    hipLaunchKernelGGL(k_sur_out_bwd<true>, dim3(nblk(n, 256)), dim3(256), 0, st, B, s->d_gout, p, Z3, dA);
  else
    hipLaunchKernelGGL(k_sur_out_bwd<false>, dim3(nblk(n, 256)), dim3(256), 0, st, B, s->d_gout, p, Z3, dA);
  // layer 3: dW3 = dZ3^T H2, db3, dH2 = dZ3 W3
  const int tw = (kSurH / 16) * (kSurH / 16);
  mm<true, false, kEpiStore>(s, tw, st, kSurH, kSurH, B, dA,
                     kSurH, H2, kSurH, g + kOffW3, kSurH, (const float*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_sur_colsum, dim3(kSurH), dim3(256), 0, st, B, kSurH, dA, (const float*)nullptr, g + kOffB3);
  const int tb = (B / 16) * (kSurH / 16);
  mm<false, false, kEpiStore>(s, tb, st, B, kSurH, kSurH, dA,
                     kSurH, p + kOffW3, kSurH, dB, kSurH, (const float*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_sur_silu_bwd, dim3(nblk(n, 256)), dim3(256), 0, st, n, Z2, dB);  // dZ2
  // layer 2
  mm<true, false, kEpiStore>(s, tw, st, kSurH, kSurH, B, dB,
                     kSurH, H1, kSurH, g + kOffW2, kSurH, (const float*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_sur_colsum, dim3(kSurH), dim3(256), 0, st, B, kSurH, dB, (const float*)nullptr, g + kOffB2);
  mm<false, false, kEpiStore>(s, tb, st, B, kSurH, kSurH, dB,
                     kSurH, p + kOffW2, kSurH, dA, kSurH, (const float*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_sur_silu_bwd, dim3(nblk(n, 256)), dim3(256), 0, st, n, Z1, dA);  // dZ1
  // layer 1: dW1 = dZ1^T X [128 x 32], db1
  const int t1 = (kSurH / 16) * (kSurIn / 16);
  mm<true, false, kEpiStore>(s, t1, st, kSurH, kSurIn, B, dA,
                     kSurH, X, kSurIn, g + kOffW1, kSurIn, (const float*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_sur_colsum, dim3(kSurH), dim3(256), 0, st, B, kSurH, dA, (const float*)nullptr, g + kOffB1);
  UPC_HIP(hipGetLastError());
  return 0;
}

void free_sur(upc_sur* s)
{
  void* ptrs[] = {s->d_p, s->d_m, s->d_v, s->d_g, s->d_act, s->d_xb, s->d_tb, s->d_pred, s->d_gout, s->d_loss,
                  s->d_step, s->d_img};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
  delete s;
}

int check_batch(int B)
{
  if (B < 32 || B > kSurMaxBatch || B % 32) {
    set_error("batch size must be a multiple of 32 in [32, 65536]");
    return UPC_EINVAL;
  }
  return 0;
}

// Host initialisation: uniform(+-sqrt(6 / (fan_in + fan_out))) weights, zero biases.
void init_params(float* p, unsigned long long seed, int net)
{
  std::vector<float> w(kSurParams, 0.f);
  auto fill = [&](int off, int out, int in, int used_in) {
    const double lim = std::sqrt(6.0 / (used_in + out));
    for (int o = 0; o < out; ++o)
      for (int i = 0; i < in; ++i) {
        if (i >= used_in) continue;
        Philox4x32 ph;
        ph.block(seed + 7919ull * net, (unsigned long long)(off + o * in + i), 1);
        const double u = philox_u01(ph.c[0], ph.c[1]);
        p[off + o * in + i] = (float)((2 * u - 1) * lim);
      }
  };
  std::memcpy(p, w.data(), sizeof(float) * kSurParams);
  fill(kOffW1, kSurH, kSurIn, kSurFeat);
  fill(kOffW2, kSurH, kSurH, kSurH);
  fill(kOffW3, kSurH, kSurH, kSurH);
  fill(kOffW4, 1, kSurH, kSurH);
}

}  // namespace

namespace upc {
// Grid form used by upc_lumi_grid(UPC_FORM_SURROGATE).
// Fraction of the error bound that the anchors around a bin must meet.
constexpr double kAnchorMargin = 0.5;

struct SurAttach {
  upc_sur* s;
  double max_err;
  int anchor_m = 16, anchor_y = 4;  // anchor spacing in M and Y bins (0: ensemble-spread gate)
  long long served = 0, fallback = 0, anchors = 0;
};

int lumi_grid_surrogate(upc_ctx* c, double* d_L, double* d_Ls, double* d_Lp, hipStream_t st)
{
  (void)d_Ls;
  (void)d_Lp;
  auto* at = static_cast<SurAttach*>(c->surrogate);
  if (!at || !at->s) return set_error("no surrogate attached (upc_ctx_set_surrogate)"), UPC_ESTATE;
  if (c->cfg.use_polarized) return set_error("the surrogate predicts the unpolarised luminosity"), UPC_EINVAL;
  upc_sur* s = at->s;
  int rc = make_grid_points(c, st);
  if (rc) return rc;
  const long n = (long)c->cfg.nm * c->cfg.ny;
  const bool anchored = at->anchor_m > 0 && at->anchor_y > 0;
  AnchorGrid ag{c->cfg.nm, c->cfg.ny, at->anchor_m, at->anchor_y, 0, 0};
  if (anchored) {
    ag.nam = anchor_count(ag.nm, ag.em);
    ag.nay = anchor_count(ag.ny, ag.ey);
  }
  const int na = ag.nam * ag.nay;
  float *X = nullptr, *pa = nullptr, *pb = nullptr;
  double *logL0 = nullptr, *Lf = nullptr, *pts = nullptr, *La = nullptr, *aerr = nullptr;
  int* idx = nullptr;
  unsigned long long* nf = nullptr;
  UPC_HIP(hipMallocAsync((void**)&X, sizeof(float) * kSurIn * n, st));
  UPC_HIP(hipMallocAsync((void**)&pa, sizeof(float) * 2 * n, st));
  UPC_HIP(hipMallocAsync((void**)&logL0, sizeof(double) * n, st));
  UPC_HIP(hipMallocAsync((void**)&idx, sizeof(int) * n, st));
  UPC_HIP(hipMallocAsync((void**)&nf, 2 * sizeof(unsigned long long), st));
  // the subset grids below rebuild (and may reallocate) the work buffer: keep the grid's (M, Y, dm, dy)
  UPC_HIP(hipMallocAsync((void**)&pts, sizeof(double) * 4 * n, st));
  UPC_HIP(hipMemcpyAsync(pts, c->work.d_points, sizeof(double) * 4 * n, hipMemcpyDeviceToDevice, st));
  pb = pa + n;
  {
    PhaseScope ph(c, st, UPC_PHASE_LUMI);
    hipLaunchKernelGGL(k_sur_features, dim3(nblk(n, 256)), dim3(256), 0, st, sur_system(c->cfg), pts, pts + 1, n,
                       4, X, logL0);
    for (int net = 0; net < kSurNets; ++net) {
      rc = infer(s, net, X, (int)n, net ? pb : pa, st);  // any row count, one launch
      if (rc) return rc;
    }
  }
  UPC_HIP(hipMemsetAsync(nf, 0, 2 * sizeof(unsigned long long), st));
  unsigned char* need_before = nullptr;
  double* Lq = nullptr;
  if (anchored) {
    // The network judged by the FP64 quadrature on the anchor lattice.  The
    // bins outside its domain go to the quadrature whatever the network
    // says, so their flux rows run in the same QAGS launch as the anchors'
    // (the anchors alone are too few integrals to fill the GPU); only the
    // bins that the gate then rejects need a second launch.
    UPC_HIP(hipMallocAsync((void**)&La, sizeof(double) * na, st));
    UPC_HIP(hipMallocAsync((void**)&aerr, sizeof(double) * na, st));
    UPC_HIP(hipMallocAsync((void**)&need_before, (size_t)n, st));
    UPC_HIP(hipMallocAsync((void**)&Lq, sizeof(double) * n, st));
    hipLaunchKernelGGL(k_sur_need_before, dim3(nblk(n, 256)), dim3(256), 0, st, ag, logL0, need_before, nf + 1);
    rc = lumi_grid_subset_fp64(c, need_before, Lq, st);
    if (rc) return rc;
    PhaseScope ph(c, st, UPC_PHASE_LUMI);
    hipLaunchKernelGGL(k_sur_anchor_vals, dim3(nblk(na, 256)), dim3(256), 0, st, ag, Lq, pts, La);
    hipLaunchKernelGGL(k_sur_anchor_err, dim3(nblk(na, 256)), dim3(256), 0, st, ag, pa, pb, logL0, s->mu,
                       s->sigma, La, aerr);
    hipLaunchKernelGGL(k_sur_gate, dim3(nblk(n, 256)), dim3(256), 0, st, ag, pa, pb, logL0, s->mu, s->sigma, aerr,
                       need_before, Lq, kAnchorMargin * at->max_err, pts, d_L, idx, nf);
  } else {
    PhaseScope ph(c, st, UPC_PHASE_LUMI);
    hipLaunchKernelGGL(k_sur_combine, dim3(nblk(n, 256)), dim3(256), 0, st, n, pa, pb, logL0, s->mu, s->sigma,
                       s->err_scale, at->max_err, pts, d_L, idx, nf, (double*)nullptr);
  }
  unsigned long long h_cnt[2] = {0, 0};  // bins rejected by the gate, bins answered before it
  UPC_HIP(hipMemcpyAsync(h_cnt, nf, sizeof(h_cnt), hipMemcpyDeviceToHost, st));
  UPC_HIP(hipStreamSynchronize(st));
  const unsigned long long h_nf = h_cnt[0];
  at->anchors = na;
  at->served = n - (long long)h_cnt[1] - (long long)h_nf;
  at->fallback = (long long)h_nf + (long long)h_cnt[1];
  if (h_nf) {
    // the fallback bins by the FP64 grid restricted to them: their flux rows
    // (shared between mirror bins as in the full grid) and their
    // contraction items only
    unsigned char* need = nullptr;
    UPC_HIP(hipMallocAsync((void**)&need, (size_t)n, st));
    UPC_HIP(hipMallocAsync((void**)&Lf, sizeof(double) * n, st));
    UPC_HIP(hipMemsetAsync(need, 0, (size_t)n, st));
    hipLaunchKernelGGL(k_sur_mark, dim3(nblk(h_nf, 256)), dim3(256), 0, st, (long)h_nf, idx, need);
    rc = lumi_grid_subset_fp64(c, need, Lf, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sur_take, dim3(nblk(h_nf, 256)), dim3(256), 0, st, (long)h_nf, idx, Lf, d_L);
    (void)hipFreeAsync(need, st);
    (void)hipFreeAsync(Lf, st);
  }
  void* bufs[] = {X, pa, logL0, idx, nf, pts, La, aerr, need_before, Lq};
  for (void* b : bufs)
    if (b) (void)hipFreeAsync(b, st);
  UPC_HIP(hipGetLastError());
  return 0;
}

void surrogate_counts(upc_ctx* c, long long* served, long long* fallback)
{
  auto* at = static_cast<SurAttach*>(c->surrogate);
  *served = at ? at->served : 0;
  *fallback = at ? at->fallback : 0;
}

void free_surrogate(upc_ctx* c)
{
  delete static_cast<SurAttach*>(c->surrogate);
  c->surrogate = nullptr;
}
}  // namespace upc

extern "C" {

int upc_sur_create(unsigned long long seed, upc_sur** out)
{
  if (!out) return set_error("null output"), UPC_EINVAL;
  auto* s = new upc_sur();
  s->seed = seed;
  auto alloc = [&]() -> int {
    UPC_HIP(hipMalloc(&s->d_p, sizeof(float) * kSurParams * kSurNets));
    UPC_HIP(hipMalloc(&s->d_m, sizeof(float) * kSurParams * kSurNets));
    UPC_HIP(hipMalloc(&s->d_v, sizeof(float) * kSurParams * kSurNets));
    UPC_HIP(hipMalloc(&s->d_g, sizeof(float) * kSurParams));
    UPC_HIP(hipMalloc(&s->d_act, sizeof(float) * 8L * kSurMaxBatch * kSurH));
    UPC_HIP(hipMalloc(&s->d_xb, sizeof(float) * (long)kSurMaxBatch * kSurIn));
    UPC_HIP(hipMalloc(&s->d_img, sizeof(unsigned short) * kInferWords));
    UPC_HIP(hipMalloc(&s->d_tb, sizeof(float) * kSurMaxBatch));
    UPC_HIP(hipMalloc(&s->d_pred, sizeof(float) * kSurMaxBatch));
    UPC_HIP(hipMalloc(&s->d_gout, sizeof(float) * kSurMaxBatch));
    UPC_HIP(hipMalloc(&s->d_step, sizeof(int)));
    UPC_HIP(hipMemset(s->d_m, 0, sizeof(float) * kSurParams * kSurNets));
    UPC_HIP(hipMemset(s->d_v, 0, sizeof(float) * kSurParams * kSurNets));
    UPC_HIP(hipMemset(s->d_step, 0, sizeof(int)));
    std::vector<float> h(kSurParams);
    for (int net = 0; net < kSurNets; ++net) {
      init_params(h.data(), seed, net);
      UPC_HIP(hipMemcpy(s->d_p + (long)net * kSurParams, h.data(), sizeof(float) * kSurParams,
                        hipMemcpyHostToDevice));
    }
    UPC_HIP(hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking));
    return 0;
  };
  if (int rc = alloc()) {
    free_sur(s);
    return rc;
  }
  *out = s;
  return 0;
}

int upc_sur_destroy(upc_sur* s)
{
  if (!s) return 0;
  (void)hipDeviceSynchronize();
  free_sur(s);
  return 0;
}

int upc_sur_param_count(int* h_n)
{
  if (!h_n) return set_error("null pointer"), UPC_EINVAL;
  *h_n = kSurParams;
  return 0;
}

int upc_sur_get_params(upc_sur* s, int net, float* h_out)
{
  if (!s || !h_out || net < 0 || net >= kSurNets) return set_error("bad arguments"), UPC_EINVAL;
  UPC_HIP(hipDeviceSynchronize());
  UPC_HIP(hipMemcpy(h_out, s->d_p + (long)net * kSurParams, sizeof(float) * kSurParams, hipMemcpyDeviceToHost));
  return 0;
}

int upc_sur_set_params(upc_sur* s, int net, const float* h_in)
{
  if (!s || !h_in || net < 0 || net >= kSurNets) return set_error("bad arguments"), UPC_EINVAL;
  UPC_HIP(hipDeviceSynchronize());
  UPC_HIP(hipMemcpy(s->d_p + (long)net * kSurParams, h_in, sizeof(float) * kSurParams, hipMemcpyHostToDevice));
  return 0;
}

int upc_sur_set_norm(upc_sur* s, double mu, double sigma)
{
  if (!s || !(sigma > 0)) return set_error("bad arguments"), UPC_EINVAL;
  s->mu = mu;
  s->sigma = sigma;
  return 0;
}

int upc_sur_set_precision(upc_sur* s, int split)
{
  if (!s || (split != 0 && split != 1)) return set_error("bad arguments"), UPC_EINVAL;
  s->split = split;
  return 0;
}

int upc_sur_set_err_scale(upc_sur* s, double scale)
{
  if (!s || !(scale > 0)) return set_error("bad arguments"), UPC_EINVAL;
  s->err_scale = scale;
  return 0;
}

int upc_sur_get_err_scale(upc_sur* s, double* h_scale)
{
  if (!s || !h_scale) return set_error("bad arguments"), UPC_EINVAL;
  *h_scale = s->err_scale;
  return 0;
}

int upc_sur_get_norm(upc_sur* s, double* h_mu, double* h_sigma)
{
  if (!s || !h_mu || !h_sigma) return set_error("bad arguments"), UPC_EINVAL;
  *h_mu = s->mu;
  *h_sigma = s->sigma;
  return 0;
}

int upc_sur_features(const upc_config* cfg, const double* d_M, const double* d_Y, long long n, float* d_X,
                     double* d_logL0, void* stream)
{
  if (!cfg || !d_M || !d_Y || !d_X || !d_logL0 || n < 0) return set_error("bad arguments"), UPC_EINVAL;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_sur_features, dim3(nblk(n, 256)), dim3(256), 0, (hipStream_t)stream, sur_system(*cfg), d_M,
                     d_Y, (long)n, 1, d_X, d_logL0);
  UPC_HIP(hipGetLastError());
  return 0;
}

int upc_sur_forward(upc_sur* s, int net, const float* d_X, int B, float* d_pred, void* stream)
{
  if (!s || net < 0 || net >= kSurNets || !d_X || !d_pred) return set_error("bad arguments"), UPC_EINVAL;
  if (int rc = check_batch(B)) return rc;
  return forward(s, net, d_X, B, nullptr, d_pred, nullptr, (hipStream_t)stream);
}

int upc_sur_grad(upc_sur* s, int net, const float* d_X, const float* d_T, int B, float* d_grad, void* stream)
{
  if (!s || net < 0 || net >= kSurNets || !d_X || !d_T || !d_grad) return set_error("bad arguments"), UPC_EINVAL;
  if (int rc = check_batch(B)) return rc;
  hipStream_t st = (hipStream_t)stream;
  int rc = forward(s, net, d_X, B, d_T, s->d_pred, nullptr, st);
  if (!rc) rc = backward(s, net, d_X, B, st);
  if (rc) return rc;
  UPC_HIP(hipMemcpyAsync(d_grad, s->d_g, sizeof(float) * kSurParams, hipMemcpyDeviceToDevice, st));
  return 0;
}

int upc_sur_train(upc_sur* s, const float* d_X, const float* d_T, long long n, int steps, int batch, double lr,
                  int use_graph, void* stream, float* h_loss)
{
  if (!s || !d_X || !d_T || n <= 0 || steps <= 0) return set_error("bad arguments"), UPC_EINVAL;
  if (int rc = check_batch(batch)) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (s->loss_cap < steps + 1) {
    if (s->d_loss) (void)hipFree(s->d_loss);
    UPC_HIP(hipMalloc(&s->d_loss, sizeof(float) * (steps + 1)));
    s->loss_cap = steps + 1;
  }
  const AdamCfg ac{(float)lr, (float)(lr * 1e-2), 0.9f, 0.999f, 1e-8f, steps};
  for (int net = 0; net < kSurNets; ++net) {
    float* p = s->d_p + (long)net * kSurParams;
    float* m = s->d_m + (long)net * kSurParams;
    float* v = s->d_v + (long)net * kSurParams;
    hipLaunchKernelGGL(k_batch_zero_f, dim3(nblk(steps + 1, 256)), dim3(256), 0, st, s->d_loss, steps + 1);
    hipLaunchKernelGGL(k_batch_zero_f, dim3(nblk(kSurParams, 256)), dim3(256), 0, st, m, kSurParams);
    hipLaunchKernelGGL(k_batch_zero_f, dim3(nblk(kSurParams, 256)), dim3(256), 0, st, v, kSurParams);
    hipLaunchKernelGGL(k_batch_zero_i, dim3(1), dim3(1), 0, st, s->d_step);
    const unsigned long long seed = s->seed * 1000003ull + 17ull * net;
    auto step_fn = [&](hipStream_t q) -> int {
      hipLaunchKernelGGL(k_sur_step_inc, dim3(1), dim3(1), 0, q, s->d_step);
      hipLaunchKernelGGL(k_sur_gather, dim3(nblk(batch, 256)), dim3(256), 0, q, batch, (long)n, seed, s->d_step,
                         d_X, d_T, s->d_xb, s->d_tb);
      int rc = forward(s, net, s->d_xb, batch, s->d_tb, s->d_pred, s->d_loss, q);
      if (!rc) rc = backward(s, net, s->d_xb, batch, q);
      if (rc) return rc;
      hipLaunchKernelGGL(k_sur_adam, dim3(nblk(kSurParams, 256)), dim3(256), 0, q, kSurParams, p, s->d_g, m, v,
                         ac, s->d_step);
      UPC_HIP(hipGetLastError());
      return 0;
    };
    if (use_graph) {
      UPC_HIP(hipStreamSynchronize(st));
      UPC_HIP(hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal));
      const int rc = step_fn(s->cap_stream);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s->cap_stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      UPC_HIP(e);
      hipGraphExec_t ex = nullptr;
      UPC_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      for (int i = 0; i < steps; ++i) {
        const hipError_t le = hipGraphLaunch(ex, st);
        if (le != hipSuccess) {
          (void)hipGraphExecDestroy(ex);
          (void)hipGraphDestroy(g);
          return hip_fail(le, "hipGraphLaunch");
        }
      }
      UPC_HIP(hipStreamSynchronize(st));
      (void)hipGraphExecDestroy(ex);
      (void)hipGraphDestroy(g);
    } else {
      for (int i = 0; i < steps; ++i)
        if (int rc = step_fn(st)) return rc;
    }
    if (h_loss) {
      std::vector<float> tmp(steps + 1);
      UPC_HIP(hipMemcpyAsync(tmp.data(), s->d_loss, sizeof(float) * (steps + 1), hipMemcpyDeviceToHost, st));
      UPC_HIP(hipStreamSynchronize(st));
      std::memcpy(h_loss + (long)net * steps, tmp.data() + 1, sizeof(float) * steps);
    }
  }
  return 0;
}

int upc_ctx_set_surrogate_anchors(upc_ctx* c, int every_m, int every_y)
{
  if (!c) return set_error("null context"), UPC_EINVAL;
  auto* at = static_cast<SurAttach*>(c->surrogate);
  if (!at) return set_error("no surrogate attached (upc_ctx_set_surrogate)"), UPC_ESTATE;
  if (every_m < 0 || every_y < 0 || (every_m == 0) != (every_y == 0))
    return set_error("anchor spacings must both be positive, or both 0"), UPC_EINVAL;
  at->anchor_m = every_m;
  at->anchor_y = every_y;
  return 0;
}

int upc_ctx_set_surrogate(upc_ctx* c, upc_sur* s, double max_err)
{
  if (!c) return set_error("null context"), UPC_EINVAL;
  if (s && !(max_err > 0)) return set_error("max_err must be positive"), UPC_EINVAL;
  free_surrogate(c);
  if (s) {
    auto* at = new SurAttach();
    at->s = s;
    at->max_err = max_err;
    c->surrogate = at;
  }
  return 0;
}

}  // extern "C"
