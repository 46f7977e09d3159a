// Samplers with the interface of upcgen's include/UpcSampler.h:37-142.  The
// cumulative tables are built on the GPU (upc_pdf_init, gsl_histogram_pdf_init
// semantics); single draws run on the host with the same counter-based
// Philox4x32-10 stream the GPU event kernels use (seed, draw index), so the
// reference's mt19937 sequences are not reproduced, only the distributions.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../include/upc_lumi.h"
#include "UpcLog.h"
#include "upc_philox.h"

namespace upc_sampler_detail
{
// gsl_histogram_pdf_init on the GPU for `rows` rows of n bins.
inline std::vector<double> pdf_init(const double* bins, int rows, int n)
{
  std::vector<double> sums((size_t)rows * (n + 1));
  double *d = nullptr, *s = nullptr;
  if (hipMalloc(&d, sizeof(double) * (size_t)rows * n) != hipSuccess ||
      hipMalloc(&s, sizeof(double) * sums.size()) != hipSuccess) {
    UPC_LOG_FATAL << "hipMalloc failed in the sampler";
    std::_Exit(-1);
  }
  (void)hipMemcpy(d, bins, sizeof(double) * (size_t)rows * n, hipMemcpyHostToDevice);
  if (upc_pdf_init(d, rows, n, s, nullptr) != 0) {
    UPC_LOG_FATAL << "Could not initialize sampler: " << upc_last_error();
    std::_Exit(-1);
  }
  (void)hipMemcpy(sums.data(), s, sizeof(double) * sums.size(), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  (void)hipFree(s);
  return sums;
}

inline uint64_t resolve_seed(uint64_t seed)
{
  if (seed == std::numeric_limits<uint64_t>::max())
    seed = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::system_clock::now().time_since_epoch())
               .count();
  return seed;
}

// gsl_histogram_pdf_sample bin search: sum[i] <= r < sum[i+1]
inline long find(const double* sum, long n, double r)
{
  if (r < sum[0] || r >= sum[n]) return -1;
  long lo = 0, hi = n;
  while (hi - lo > 1) {
    const long mid = (lo + hi) / 2;
    if (r >= sum[mid]) lo = mid;
    else hi = mid;
  }
  return lo;
}
}  // namespace upc_sampler_detail

class UpcSampler1D
{
 public:
  UpcSampler1D(const std::vector<double>& dist, const std::vector<double>& binEdges,
               uint64_t seed = std::numeric_limits<uint64_t>::max())
      : edges(binEdges), seed_(upc_sampler_detail::resolve_seed(seed))
  {
    sum = upc_sampler_detail::pdf_init(dist.data(), 1, (int)dist.size());
  }
  // from a cumulative table already built (n + 1 entries)
  UpcSampler1D(const double* cdf, int n, const std::vector<double>& binEdges, uint64_t seed)
      : sum(cdf, cdf + n + 1), edges(binEdges), seed_(upc_sampler_detail::resolve_seed(seed))
  {
  }

  double operator()()
  {
    upc::Philox4x32 ph;
    ph.block(seed_, draw_++, 0);
    double r = upc::philox_u01(ph.c[0], ph.c[1]);
    if (r == 1.0) r = 0.0;
    const long n = (long)sum.size() - 1;
    const long i = upc_sampler_detail::find(sum.data(), n, r);
    if (i < 0) return 0.;  // GSL_EDOM in the reference
    const double delta = (r - sum[i]) / (sum[i + 1] - sum[i]);
    return edges[i] + delta * (edges[i + 1] - edges[i]);
  }

  std::vector<double> sum;    // cumulative pdf
  std::vector<double> edges;  // bin edges

 private:
  uint64_t seed_;
  uint64_t draw_{0};
};

class UpcSampler2D
{
 public:
  // dist[ix][iy], row-major as the reference fills the GSL 2-D histogram
  UpcSampler2D(const std::vector<std::vector<double>>& dist, const std::vector<double>& binEdgesX,
               const std::vector<double>& binEdgesY, uint64_t seed = std::numeric_limits<uint64_t>::max())
      : nBinsX((int)dist.size()), nBinsY((int)dist[0].size()), edgesX(binEdgesX), edgesY(binEdgesY),
        seed_(upc_sampler_detail::resolve_seed(seed))
  {
    std::vector<double> flat((size_t)nBinsX * nBinsY);
    for (int i = 0; i < nBinsX; ++i)
      for (int j = 0; j < nBinsY; ++j) flat[(size_t)i * nBinsY + j] = dist[i][j];
    sum = upc_sampler_detail::pdf_init(flat.data(), 1, (int)flat.size());
  }

  void operator()(double& x, double& y)
  {
    upc::Philox4x32 ph;
    ph.block(seed_, draw_++, 0);
    double r1 = upc::philox_u01(ph.c[0], ph.c[1]);
    double r2 = upc::philox_u01(ph.c[2], ph.c[3]);
    if (r2 == 1.0) r2 = 0.0;
    if (r1 == 1.0) r1 = 0.0;
    const long k = upc_sampler_detail::find(sum.data(), (long)nBinsX * nBinsY, r1);
    if (k < 0) {
      x = y = 0.;
      return;
    }
    const long i = k / nBinsY, j = k - i * nBinsY;
    const double delta = (r1 - sum[k]) / (sum[k + 1] - sum[k]);
    x = edgesX[i] + delta * (edgesX[i + 1] - edgesX[i]);
    y = edgesY[j] + r2 * (edgesY[j + 1] - edgesY[j]);
  }

  // from 0 to nBinsX; truncation before the division as in the reference
  int getBinX(double x) { return int(nBinsX * (x - edgesX.front())) / (edgesX.back() - edgesX.front()); }
  int getBinY(double y) { return int(nBinsY * (y - edgesY.front())) / (edgesY.back() - edgesY.front()); }

  int nBinsX;
  int nBinsY;
  std::vector<double> edgesX;
  std::vector<double> edgesY;
  std::vector<double> sum;

 private:
  uint64_t seed_;
  uint64_t draw_{0};
};
