// Elementary and nuclear cross-section tables, sampler CDFs and event draws.
//
//   UpcTwoPhotonDilep::calcCrossSection*        src/UpcTwoPhotonDilep.cpp:29-134
//   UpcCrossSection::fillCrossSectionZM         src/UpcCrossSection.cpp:337-362
//   UpcCrossSection::calcNucCrossSectionYM      src/UpcCrossSection.cpp:594-698
//   UpcSampler1D/2D (gsl_histogram*_pdf)        include/UpcSampler.h:37-142
//   UpcGenerator::generateEvent (Y, M, z draw)  src/UpcGenerator.cpp:735-757
#include "upc_internal.h"
#include "upc_philox.h"

namespace upc {

__device__ __forceinline__ double lep_mass(int pdg)
{
  return pdg == 11 ? kMEl : (pdg == 13 ? kMMu : (pdg == 15 ? kMTau : 0.));
}

// src/UpcTwoPhotonDilep.cpp:46-61  [nb]
__device__ double dilep_cs_m(double mPart, double aLep, double m)
{
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
  cs *= 4 * kHc * kHc * 1e7 * kAlpha * kAlpha * kPi / s;
  return cs;
}

// src/UpcTwoPhotonDilep.cpp:63-83  [GeV^-2]
__device__ double dilep_cs_zm(double mPart, double aLep, double z, double m)
{
  const double s = m * m;
  const double k = sqrt(s) / 2.;
  const double p = sqrt(k * k - mPart * mPart);
  const double norm = 2 * kPi * kAlpha * kAlpha / s * p / k;
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

// src/UpcTwoPhotonDilep.cpp:85-95  [fm^2]
__device__ double dilep_cs_m_pol_s(double mPart, double m)
{
  const double r = 2 * mPart / m;
  if (r > 1) return 0;
  return 4 * kPi * kAlpha * kAlpha * kHc * kHc / m / m *
         ((1 + r * r - 3. / 4. * r * r * r * r) * 2 * log(1 / r + sqrt(1 / r / r - 1)) -
          (1 + 3. / 2. * r * r) * sqrt(1 - r * r));
}

// src/UpcTwoPhotonDilep.cpp:111-121
__device__ double dilep_cs_m_pol_ps(double mPart, double m)
{
  const double r = 2 * mPart / m;
  if (r > 1) return 0;
  return 4 * kPi * kAlpha * kAlpha * kHc * kHc / m / m *
         ((1 + r * r - 1. / 4. * r * r * r * r) * 2 * log(1 / r + sqrt(1 / r / r - 1)) -
          (1 + 1. / 2. * r * r) * sqrt(1 - r * r));
}

// src/UpcTwoPhotonDilep.cpp:97-109
__device__ double dilep_cs_zm_pol_s(double mPart, double z, double m)
{
  const double mLep2 = mPart * mPart;
  const double m2 = m * m;
  const double z2 = z * z;
  double cs_s = 2 * kPi * kAlpha * kAlpha;
  cs_s *= m2 - 4 * mLep2;
  cs_s *= sqrt(m2 - 4 * mLep2);
  cs_s *= (4 * mLep2 * (3 - 2 * z2 + z2 * z2)) + m2 * (1 - z2 * z2);
  cs_s /= m2 * m * (m2 * (1 - z2) + 4 * mLep2 * z2) * (m2 * (1 - z2) + 4 * mLep2 * z2);
  return cs_s;
}

// src/UpcTwoPhotonDilep.cpp:123-133
__device__ double dilep_cs_zm_pol_ps(double mPart, double z, double m)
{
  const double mLep2 = mPart * mPart;
  const double m2 = m * m;
  const double z2 = z * z;
  double cs_s = 2 * kPi * kAlpha * kAlpha;
  cs_s *= sqrt(m2 - 4 * mLep2);
  cs_s *= m2 * m2 * (1 - z2 * z2) + 8 * m2 * mLep2 * (1 - z2 + z2 * z2) -
          16 * mLep2 * mLep2 * (1 - z2) * (1 - z2);
  cs_s /= m2 * m * (m2 * (1 - z2) + 4 * mLep2 * z2) * (m2 * (1 - z2) + 4 * mLep2 * z2);
  return cs_s;
}

// fillCrossSectionZM: out[im][iz] = cs(z, m) * hc^2 1e7 / dm at lower edges
__global__ void k_cs_zm(int flag, int pdg, double aLep, int nm, double mmin, double mmax, int nz, double zmin,
                        double zmax, double* __restrict__ out)
{
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)nm * nz) return;
  const int im = (int)(gid / nz), iz = (int)(gid % nz);
  const double scalingFactor = kHc * kHc * 1e7;
  const double dm = (mmax - mmin) / nm;
  const double dz = (zmax - zmin) / nz;
  const double m = mmin + dm * im;
  const double z = zmin + dz * iz;
  const double mp = lep_mass(pdg);
  double cs = 0;
  if (flag == 0) cs = dilep_cs_zm(mp, aLep, z, m);
  if (flag == 1) cs = dilep_cs_zm_pol_s(mp, z, m);
  if (flag == 2) cs = dilep_cs_zm_pol_ps(mp, z, m);
  out[gid] = cs * scalingFactor / dm;
}

__global__ void k_cs_m(int flag, int pdg, double aLep, int nm, double mmin, double mmax, double* __restrict__ out)
{
  const int im = blockIdx.x * blockDim.x + threadIdx.x;
  if (im >= nm) return;
  const double dm = (mmax - mmin) / nm;
  const double m = mmin + dm * im;
  const double mp = lep_mass(pdg);
  out[im] = flag == 0 ? dilep_cs_m(mp, aLep, m) : (flag == 1 ? dilep_cs_m_pol_s(mp, m) : dilep_cs_m_pol_ps(mp, m));
}

__global__ void k_elem_cs(int flag, int pdg, double aLep, const double* __restrict__ m, const double* __restrict__ z,
                          int n, double* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double mp = lep_mass(pdg);
  double v = 0;
  switch (flag) {
    case 0: v = dilep_cs_m(mp, aLep, m[i]); break;
    case 1: v = dilep_cs_m_pol_s(mp, m[i]); break;
    case 2: v = dilep_cs_m_pol_ps(mp, m[i]); break;
    case 3: v = dilep_cs_zm(mp, aLep, z[i], m[i]); break;
    case 4: v = dilep_cs_zm_pol_s(mp, z[i], m[i]); break;
    default: v = dilep_cs_zm_pol_ps(mp, z[i], m[i]); break;
  }
  out[i] = v;
}

// calcNucCrossSectionYM: cs[iy][im] from L[im][iy]; per-block partial totals.
template <bool POL>
__global__ __launch_bounds__(256) void k_nuc_cs_ym(int pdg, double aLep, int nm, double mmin, double mmax, int ny,
                                                   const double* __restrict__ L, const double* __restrict__ Ls,
                                                   const double* __restrict__ Lp, double* __restrict__ cs,
                                                   double* __restrict__ ratio, double* __restrict__ partial)
{
  __shared__ double red[256];
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (gid < (long)nm * ny) {
    const int im = (int)(gid / ny), iy = (int)(gid % ny);
    const double dm = (mmax - mmin) / nm;
    const double m = mmin + dm * im;
    const double mp = lep_mass(pdg);
    const long oi = (long)iy * nm + im;
    if (!POL) {
      v = dilep_cs_m(mp, aLep, m) * L[gid];
    } else {
      const double cs_s = dilep_cs_m_pol_s(mp, m);
      const double cs_p = dilep_cs_m_pol_ps(mp, m);
      const double nuccs_s = Ls[gid] * cs_s;
      const double nuccs_p = Lp[gid] * cs_p;
      const double nuccs = nuccs_s + nuccs_p;
      v = nuccs * 1e7;
      ratio[oi] = nuccs_s / nuccs_p;
    }
    cs[oi] = v;
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum(const double* __restrict__ x, int n, double* __restrict__ out)
{
  __shared__ double red[256];
  double v = 0;
  for (int i = threadIdx.x; i < n; i += 256) v += x[i];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// gsl_histogram_pdf_init per row: sum[0] = 0, sum[i+1] = sum[i] + (bin_i / mean) / n.
// One 256-thread block per row; the row is scanned in contiguous per-thread
// chunks, the mean as the plain sum over n.
__global__ __launch_bounds__(256) void k_pdf_init(const double* __restrict__ bins, int n,
                                                  double* __restrict__ sum, int* __restrict__ bad)
{
  __shared__ double part[256];
  __shared__ double mean_s;
  const double* row = bins + (long)blockIdx.x * n;
  double* srow = sum + (long)blockIdx.x * (n + 1);
  const int t = threadIdx.x;
  const int per = (n + 255) / 256;
  const int s = t * per, e = min(n, s + per);
  double acc = 0;
  int neg = 0;
  for (int i = s; i < e; ++i) {
    acc += row[i];
    neg |= row[i] < 0;
  }
  if (neg) atomicOr(bad, 1);
  part[t] = acc;
  __syncthreads();
  if (t == 0) {
    double tot = 0;
    for (int i = 0; i < 256; ++i) tot += part[i];
    mean_s = tot / n;
  }
  __syncthreads();
  const double mean = mean_s;
  acc = 0;
  for (int i = s; i < e; ++i) acc += (row[i] / mean) / n;
  __syncthreads();
  part[t] = acc;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const double v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  double base = t ? part[t - 1] : 0;
  if (t == 0) srow[0] = 0;
  for (int i = s; i < e; ++i) {
    base += (row[i] / mean) / n;
    srow[i + 1] = base;
  }
}

// The same table for one long row (the 2-D (Y, M) pdf) over many blocks:
// segment sums -> mean -> scaled segment sums -> per-segment offsets and the
// prefix inside each segment.  Every block re-derives the mean and its offset
// from the segment sums in one fixed order, so all blocks agree.
constexpr int kPdfSeg = 2048;  // bins per block

__global__ __launch_bounds__(256) void k_pdf_seg_sum(const double* __restrict__ row, int n,
                                                     double* __restrict__ seg, int* __restrict__ bad)
{
  __shared__ double part[256];
  const int s0 = blockIdx.x * kPdfSeg, s1 = min(n, s0 + kPdfSeg);
  double acc = 0;
  int neg = 0;
  for (int i = s0 + threadIdx.x; i < s1; i += 256) {
    acc += row[i];
    neg |= row[i] < 0;
  }
  if (neg) atomicOr(bad, 1);
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) seg[blockIdx.x] = part[0];
}

__device__ double pdf_mean(const double* __restrict__ seg, int nseg, int n)
{
  double tot = 0;
  for (int b = 0; b < nseg; ++b) tot += seg[b];
  return tot / n;
}

__global__ __launch_bounds__(256) void k_pdf_seg_scaled(const double* __restrict__ row, int n,
                                                        const double* __restrict__ seg, int nseg,
                                                        double* __restrict__ seg_scaled)
{
  __shared__ double part[256];
  __shared__ double mean_s;
  if (threadIdx.x == 0) mean_s = pdf_mean(seg, nseg, n);
  __syncthreads();
  const double mean = mean_s;
  const int s0 = blockIdx.x * kPdfSeg, s1 = min(n, s0 + kPdfSeg);
  const int per = kPdfSeg / 256;
  const int a = s0 + threadIdx.x * per, e = min(s1, a + per);
  double acc = 0;
  for (int i = a; i < e; ++i) acc += (row[i] / mean) / n;
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) seg_scaled[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(256) void k_pdf_seg_apply(const double* __restrict__ row, int n,
                                                       const double* __restrict__ seg, int nseg,
                                                       const double* __restrict__ seg_scaled,
                                                       double* __restrict__ sum)
{
  __shared__ double part[256];
  __shared__ double mean_s, base_s;
  if (threadIdx.x == 0) {
    mean_s = pdf_mean(seg, nseg, n);
    double base = 0;
    for (int b = 0; b < (int)blockIdx.x; ++b) base += seg_scaled[b];
    base_s = base;
    if (blockIdx.x == 0) sum[0] = 0;
  }
  __syncthreads();
  const double mean = mean_s;
  const int s0 = blockIdx.x * kPdfSeg, s1 = min(n, s0 + kPdfSeg);
  const int per = kPdfSeg / 256;
  const int a = s0 + threadIdx.x * per, e = min(s1, a + per);
  double acc = 0;
  for (int i = a; i < e; ++i) acc += (row[i] / mean) / n;
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const double v = threadIdx.x >= d ? part[threadIdx.x - d] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  double base = base_s + (threadIdx.x ? part[threadIdx.x - 1] : 0);
  for (int i = a; i < e; ++i) {
    base += (row[i] / mean) / n;
    sum[i + 1] = base;
  }
}

// gsl_histogram_pdf_sample on a prefix table: sum[i] <= r < sum[i+1].
__device__ __forceinline__ long pdf_find(const double* __restrict__ sum, long n, double r)
{
  if (r < sum[0] || r >= sum[n]) return -1;
  long lo = 0, hi = n;
  while (hi - lo > 1) {
    const long mid = (lo + hi) >> 1;
    if (r >= sum[mid]) lo = mid;
    else hi = mid;
  }
  return lo;
}

// One event per thread: (Y, M) from the 2-D pdf over [ny][nm] (x = Y, y = M),
// then z from the z pdf of the M row picked by UpcSampler2D::getBinY
// (truncation before division kept), as src/UpcGenerator.cpp:735-757.
__global__ void k_sample_events(const double* __restrict__ sum_ym, const double* __restrict__ sum_z,
                                const double* __restrict__ sum_zs, const double* __restrict__ sum_zp,
                                const double* __restrict__ ratio, int pol, int nm, double mmin, double mmax,
                                int ny, double ymin, double ymax, int nz, double zmin, double zmax,
                                unsigned long long seed, unsigned long long offset, long long n,
                                double* __restrict__ oy, double* __restrict__ om, double* __restrict__ oz)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Philox4x32 ph;
  const unsigned long long ctr = offset + (unsigned long long)i;
  ph.block(seed, ctr, 0);
  double r1 = philox_u01(ph.c[0], ph.c[1]);
  double r2 = philox_u01(ph.c[2], ph.c[3]);
  ph.block(seed, ctr, 1);
  double r3 = philox_u01(ph.c[0], ph.c[1]);
  const double r4 = philox_u01(ph.c[2], ph.c[3]);
  const long nym = (long)ny * nm;
  if (r1 == 1.0) r1 = 0.0;
  if (r2 == 1.0) r2 = 0.0;
  const long kbin = pdf_find(sum_ym, nym, r1);
  if (kbin < 0) {
    oy[i] = om[i] = oz[i] = __builtin_nan("");
    return;
  }
  const long iy = kbin / nm, jm = kbin - iy * nm;
  const double dy = (ymax - ymin) / ny, dm = (mmax - mmin) / nm;
  const double y0 = ymin + dy * iy, y1 = ymin + dy * (iy + 1);
  const double m0 = mmin + dm * jm, m1 = mmin + dm * (jm + 1);
  const double delta = (r1 - sum_ym[kbin]) / (sum_ym[kbin + 1] - sum_ym[kbin]);
  const double y = y0 + delta * (y1 - y0);
  const double m = m0 + r2 * (m1 - m0);
  // getBinX / getBinY of UpcSampler2D (include/UpcSampler.h:124-133)
  const double yl = ymin + dy * ny;
  const double ml = mmin + dm * nm;
  const int ybin = (int)((double)(int)(ny * (y - ymin)) / (yl - ymin));
  int mbin = (int)((double)(int)(nm * (m - mmin)) / (ml - mmin));
  if (mbin >= nm) mbin = nm - 1;
  const double* zt = sum_z;
  if (pol) {
    const int yb = ybin >= ny ? ny - 1 : ybin;
    const double frac = ratio[(long)yb * nm + mbin];
    zt = r4 < frac ? sum_zs : sum_zp;
  }
  zt += (long)mbin * (nz + 1);
  if (r3 == 1.0) r3 = 0.0;
  const long zb = pdf_find(zt, nz, r3);
  double z = __builtin_nan("");
  if (zb >= 0) {
    const double dz = (zmax - zmin) / nz;
    const double z0 = zmin + dz * zb, z1 = zmin + dz * (zb + 1);
    z = z0 + ((r3 - zt[zb]) / (zt[zb + 1] - zt[zb])) * (z1 - z0);
  }
  oy[i] = y;
  om[i] = m;
  oz[i] = z;
}

// ---------------------------------------------------------------------------
static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

int launch_cs_zm(upc_ctx* c, double zmin, double zmax, int nz, double mmin, double mmax, int nm, int flag,
                 double* d_zm, hipStream_t s)
{
  const upc_config& g = c->cfg;
  PhaseScope ph(c, s, UPC_PHASE_XSEC);
  hipLaunchKernelGGL(k_cs_zm, dim3(nblk((long)nm * nz, 256)), dim3(256), 0, s, flag, g.proc_id, g.lep_a, nm, mmin,
                     mmax, nz, zmin, zmax, d_zm);
  UPC_HIP(hipGetLastError());
  return 0;
}

int launch_elem_cs(upc_ctx* c, int flag, const double* d_m, const double* d_z, int n, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_elem_cs, dim3(nblk(n, 256)), dim3(256), 0, s, flag, c->cfg.proc_id, c->cfg.lep_a, d_m, d_z, n,
                     d_out);
  UPC_HIP(hipGetLastError());
  return 0;
}

int launch_cs_m(upc_ctx* c, int flag, double* d_cs, hipStream_t s)
{
  const upc_config& g = c->cfg;
  hipLaunchKernelGGL(k_cs_m, dim3(nblk(g.nm, 256)), dim3(256), 0, s, flag, g.proc_id, g.lep_a, g.nm, g.mmin,
                     g.mmax, d_cs);
  UPC_HIP(hipGetLastError());
  return 0;
}

int launch_nuc_cs_ym(upc_ctx* c, const double* d_L, const double* d_Ls, const double* d_Lp, double* d_cs,
                     double* d_ratio, double* h_tot, hipStream_t s)
{
  const upc_config& g = c->cfg;
  const long n = (long)g.nm * g.ny;
  const int nb = nblk(n, 256);
  double* d_part = nullptr;
  PhaseScope* ph = new PhaseScope(c, s, UPC_PHASE_XSEC);
  UPC_HIP(hipMallocAsync((void**)&d_part, sizeof(double) * (nb + 1), s));
  if (g.use_polarized) {
    hipLaunchKernelGGL(k_nuc_cs_ym<true>, dim3(nb), dim3(256), 0, s, g.proc_id, g.lep_a, g.nm, g.mmin, g.mmax,
                       g.ny, d_L, d_Ls, d_Lp, d_cs, d_ratio, d_part);
  } else {
    hipLaunchKernelGGL(k_nuc_cs_ym<false>, dim3(nb), dim3(256), 0, s, g.proc_id, g.lep_a, g.nm, g.mmin, g.mmax,
                       g.ny, d_L, d_Ls, d_Lp, d_cs, d_ratio, d_part);
  }
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, d_part, nb, d_part + nb);
  delete ph;
  UPC_HIP(hipGetLastError());
  double tot = 0;
  UPC_HIP(hipMemcpyAsync(&tot, d_part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipFreeAsync(d_part, s));
  UPC_HIP(hipStreamSynchronize(s));
  if (h_tot) {
    *h_tot += tot;
    *h_tot *= 1e-6;
  }
  return 0;
}

int launch_pdf_init(const double* d_bins, int nrows, int n, double* d_sum, hipStream_t s)
{
  if (nrows <= 0 || n <= 0) return 0;
  int* d_bad = nullptr;
  UPC_HIP(hipMallocAsync((void**)&d_bad, sizeof(int), s));
  UPC_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
  const int nseg = (n + kPdfSeg - 1) / kPdfSeg;
  double* d_seg = nullptr;
  if (nrows == 1 && nseg > 1) {
    UPC_HIP(hipMallocAsync((void**)&d_seg, sizeof(double) * 2 * nseg, s));
    hipLaunchKernelGGL(k_pdf_seg_sum, dim3(nseg), dim3(256), 0, s, d_bins, n, d_seg, d_bad);
    hipLaunchKernelGGL(k_pdf_seg_scaled, dim3(nseg), dim3(256), 0, s, d_bins, n, d_seg, nseg, d_seg + nseg);
    hipLaunchKernelGGL(k_pdf_seg_apply, dim3(nseg), dim3(256), 0, s, d_bins, n, d_seg, nseg, d_seg + nseg,
                       d_sum);
    UPC_HIP(hipFreeAsync(d_seg, s));
  } else {
    hipLaunchKernelGGL(k_pdf_init, dim3(nrows), dim3(256), 0, s, d_bins, n, d_sum, d_bad);
  }
  UPC_HIP(hipGetLastError());
  int bad = 0;
  UPC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
  UPC_HIP(hipFreeAsync(d_bad, s));
  UPC_HIP(hipStreamSynchronize(s));
  if (bad) {
    set_error("histogram bins must be non-negative to compute a probability distribution");
    return UPC_EINVAL;
  }
  return 0;
}

int launch_sample_events(upc_ctx* c, const double* d_sum_ym, const double* d_sum_z, const double* d_sum_zs,
                         const double* d_sum_zp, const double* d_ratio, unsigned long long seed,
                         unsigned long long offset, long long n, double* d_y, double* d_m, double* d_z,
                         hipStream_t s)
{
  if (n <= 0) return 0;
  const upc_config& g = c->cfg;
  const int pol = g.use_polarized;
  if (pol ? (!d_sum_zs || !d_sum_zp || !d_ratio) : !d_sum_z) {
    set_error("missing z sampler tables");
    return UPC_EINVAL;
  }
  hipLaunchKernelGGL(k_sample_events, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_sum_ym, d_sum_z,
                     d_sum_zs, d_sum_zp, d_ratio, pol, g.nm, g.mmin, g.mmax, g.ny, g.ymin, g.ymax, g.nz, g.zmin,
                     g.zmax, seed, offset, n, d_y, d_m, d_z);
  UPC_HIP(hipGetLastError());
  return 0;
}

}  // namespace upc
