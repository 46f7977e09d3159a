// Flux-row layout and the per-node phi sum shared by the FP64 and MFMA
// luminosity forms.
#pragma once

#include "upc_internal.h"

namespace upc {

struct FluxParams {
  double bmin, R, g1, g2, factor, twoR;
  int A, isPoint;
  int mirror_ny;  // > 0: symmetric grid with ny Y bins (k2 rows shared with mirror bins)
};

// Row holding k2 of point p.
__device__ __forceinline__ int k2_row(int p, int npts, int mirror_ny)
{
  if (mirror_ny <= 0) return npts + p;
  const int im = p / mirror_ny, iy = p - im * mirror_ny;
  return iy ? im * mirror_ny + (mirror_ny - iy) : npts + im;
}

__device__ __forceinline__ double row_node(double bmin, double logd, int j, double* bl, double* bh)
{
  *bl = bmin * exp(j * logd);
  *bh = bmin * exp((j + 1.) * logd);
  return (*bh + *bl) / 2.;
}

__device__ __forceinline__ double node_b2(double b1, double b2, double c)
{
  return fma(2.0 * b1 * b2, c, fma(b1, b1, b2 * b2));
}

// The phi sum of one node pair: sum_k w_k P(min(b, 20)) G(b) (G = 1 for
// b >= 20); polarised: cos^2 and sin^2 weighted sums.
template <bool POL>
__device__ __forceinline__ void node_k(double b1, double b2, const double* cosk, const double* w5,
                                       int mode, double p20, const double4* gaa, const double4* __restrict__ bk,
                                       double* ks, double* kp)
{
  const double db = 20. / (kNB - 1);
  double s = 0.0, p = 0.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double c = cosk[k];
    const double b = sqrt(node_b2(b1, b2, POL ? -c : c));
    double breakup = 1., gaa_v = 1.;
    if (b < 20.) {
      if (mode != 1) breakup = spline_eval(bk, kBkMin, kDBk, 1.0 / kDBk, kNBkUsed, b);
      gaa_v = spline_eval(gaa, 0.0, db, 1.0 / db, kNB, b);
    } else if (mode != 1) {
      breakup = p20;
    }
    if (POL) {  // w5[5 + k] = sin^2(phi_k)
      s += breakup * gaa_v * w5[k] * c * c;
      p += breakup * gaa_v * w5[k] * w5[5 + k];
    } else {
      s += breakup * gaa_v * w5[k];
    }
  }
  *ks = s;
  *kp = p;
}

inline FluxParams flux_params(const upc_ctx* c, int mirror_ny = 0)
{
  FluxParams fp;
  fp.mirror_ny = mirror_ny;
  fp.R = c->cfg.R;
  fp.bmin = c->cfg.flux_point ? c->cfg.R : 0.05 * c->cfg.R;
  fp.g1 = c->g1;
  fp.g2 = c->g2;
  fp.factor = c->factor;
  fp.twoR = 2. * c->cfg.R;
  fp.A = c->cfg.A;
  fp.isPoint = c->cfg.flux_point;
  return fp;
}

// Builds the flux rows of the points in the work buffer (d_pts == nullptr)
// or of the npts (M, Y) pairs at d_pts ([M..., Y...]): point-charge nodes
// directly, form-factor nodes by the QAGS task kernels.  *fp receives the
// row parameters (mirror_ny set when the grid's Y range is symmetric).
int prepare_flux(upc_ctx* c, const double* d_pts, int npts, FluxParams* fp, hipStream_t s);
// Copies the QAGS counters of the last prepare_flux into the context.
int read_flux_stats(upc_ctx* c, hipStream_t s);

}  // namespace upc
