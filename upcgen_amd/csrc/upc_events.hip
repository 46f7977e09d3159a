// Event generation for the dilepton processes on the GPU:
//   UpcGenerator::generateEvent      src/UpcGenerator.cpp:715-790
//   UpcCrossSection::getPairMomentum src/UpcCrossSection.cpp:1053-1074
//   UpcCrossSection::getPhotonPt     src/UpcCrossSection.cpp:1021-1051
//   UpcGenerator::pairProduction     src/UpcGenerator.cpp:388-423
//   UpcGenerator::checkKinCuts       src/UpcGenerator.cpp:563-587
//
// getPhotonPt draws the photon transverse momentum from a 5000-bin histogram
// over (0, 6 hc / R] with contents F(pt^2 + E^2/g^2)^2 pt^3 / (4 pi^2 (pt^2 +
// E^2/g^2)^2), rebuilt per MeV of photon energy E.  The energy enters only
// through eps = E^2/g^2, so the histograms are tabulated once per context on
// a log grid of eps and interpolated linearly between neighbours; the draw
// inside the chosen bin is uniform, as TH1::GetRandom does.
#include "upc_internal.h"
#include "upc_philox.h"

namespace upc {

constexpr int kPtBins = 5000;
constexpr int kPtEps = 2048;          // eps grid points
constexpr double kPtEpsLo = -13.0;    // log10 eps range; below it the shape no longer changes
constexpr double kPtEpsHi = 0.4;      // 10^0.4 ~ 2.5 > Q2max: the form factor table ends at Q2 = 2

struct PtTables {
  double* cdf = nullptr;  // [kPtEps][kPtBins + 1]
  double gtot = 0;
  double ptmax = 0;
};

__device__ __forceinline__ double ff_clamped(const double4* __restrict__ ff, double q2)
{
  // beyond the table the reference's spline raises GSL_EDOM; here it is clamped
  const double v = q2 < kQ2min ? kQ2min : (q2 > kQ2max - kDQ2 ? kQ2max - kDQ2 : q2);
  return spline_eval(ff, kQ2min, kDQ2, 1.0 / kDQ2, kNQ2, v);
}

// One block per eps value: bin contents, then the normalised cumulative sum.
__global__ __launch_bounds__(256) void k_pt_tables(const double4* __restrict__ ff, double ptmax, double* __restrict__ cdf)
{
  __shared__ double part[256];
  const int e = blockIdx.x;
  const double eps = pow(10.0, kPtEpsLo + (kPtEpsHi - kPtEpsLo) * e / (kPtEps - 1));
  double* row = cdf + (size_t)e * (kPtBins + 1);
  constexpr double pi2x4 = 4 * kPi * kPi;
  const int t = threadIdx.x;
  const int per = (kPtBins + 255) / 256;
  const int s = t * per, en = min(kPtBins, s + per);
  double acc = 0;
  for (int b = s; b < en; ++b) {
    const double pt = ptmax / kPtBins * (b + 1);  // upper bin edge, as the reference fills bin b+1
    const double arg = pt * pt + eps;
    const double f = ff_clamped(ff, arg);
    const double prob = (f * f) * pt * pt * pt / (pi2x4 * arg * arg);
    row[b + 1] = prob;
    acc += prob;
  }
  part[t] = acc;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const double v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  const double total = part[255];
  double base = t ? part[t - 1] : 0;
  for (int b = s; b < en; ++b) {
    base += row[b + 1];
    row[b + 1] = base / total;
  }
  if (t == 0) row[0] = 0;
}

__device__ double photon_pt(const double* __restrict__ cdf, double ptmax, double E, double gtot, double r)
{
  const double eps = (E * E) / (gtot * gtot);
  const double le = eps > 0 ? log10(eps) : kPtEpsLo;
  double u = (le - kPtEpsLo) / (kPtEpsHi - kPtEpsLo) * (kPtEps - 1);
  u = u < 0 ? 0 : (u > kPtEps - 1 ? kPtEps - 1 : u);
  int i0 = (int)u;
  if (i0 > kPtEps - 2) i0 = kPtEps - 2;
  const double w = u - i0;
  const double* c0 = cdf + (size_t)i0 * (kPtBins + 1);
  const double* c1 = c0 + (kPtBins + 1);
  auto C = [&](int k) { return (1 - w) * c0[k] + w * c1[k]; };
  int lo = 0, hi = kPtBins;  // C(lo) <= r < C(hi)
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (C(mid) <= r) lo = mid;
    else hi = mid;
  }
  const double cl = C(lo), ch = C(lo + 1);
  const double frac = ch > cl ? (r - cl) / (ch - cl) : 0.5;
  return ptmax / kPtBins * (lo + frac);
}

__device__ __forceinline__ long pdf_find2(const double* __restrict__ sum, long n, double r)
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

struct EvGrid {
  int nm, ny, nz, pol, pdg;
  double mmin, mmax, ymin, ymax, zmin, zmax, mlep;
};

// TLorentzVector::Boost(bx, by, bz)
__device__ __forceinline__ void boost(double bx, double by, double bz, double* p)
{
  const double b2 = bx * bx + by * by + bz * bz;
  if (b2 <= 0) return;
  const double gamma = 1.0 / sqrt(1.0 - b2);
  const double bp = bx * p[0] + by * p[1] + bz * p[2];
  const double gamma2 = (gamma - 1.0) / b2;
  const double f = gamma2 * bp + gamma * p[3];
  p[0] += f * bx;
  p[1] += f * by;
  p[2] += f * bz;
  p[3] = gamma * (p[3] + bp);
}

__device__ __forceinline__ bool pass_cuts(const double* p, const upc_event_opts& o)
{
  const double pt = sqrt(p[0] * p[0] + p[1] * p[1]);
  if (o.do_pt_cut && pt < o.pt_min) return false;
  if (o.do_eta_cut) {
    const double eta = pt > 0 ? asinh(p[2] / pt) : (p[2] >= 0 ? 1e10 : -1e10);
    if (eta < o.eta_min || eta > o.eta_max) return false;
  }
  return true;
}

__global__ void k_generate_pairs(EvGrid g, upc_event_opts o, const double* __restrict__ sum_ym,
                                 const double* __restrict__ sum_z, const double* __restrict__ sum_zs,
                                 const double* __restrict__ sum_zp, const double* __restrict__ ratio,
                                 const double* __restrict__ ptcdf, double ptmax, double gtot,
                                 unsigned long long seed, unsigned long long offset, long long n,
                                 double* __restrict__ out_p, int* __restrict__ out_pdg, int* __restrict__ out_pass)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long ev = offset + (unsigned long long)i;
  Philox4x32 ph;
  double r[12];
  for (uint32_t b = 0; b < 6; ++b) {
    ph.block(seed, ev, b);
    r[2 * b] = philox_u01(ph.c[0], ph.c[1]);
    r[2 * b + 1] = philox_u01(ph.c[2], ph.c[3]);
  }
  double* P = out_p + 8 * i;
  out_pass[i] = 0;
  // (Y, M) from the 2-D pdf (UpcSampler2D, include/UpcSampler.h:118-121)
  const long nym = (long)g.ny * g.nm;
  const long kbin = pdf_find2(sum_ym, nym, r[0]);
  if (kbin < 0) return;
  const long iy = kbin / g.nm, jm = kbin - iy * g.nm;
  const double dy = (g.ymax - g.ymin) / g.ny, dm = (g.mmax - g.mmin) / g.nm;
  const double y0 = g.ymin + dy * iy, y1 = g.ymin + dy * (iy + 1);
  const double m0 = g.mmin + dm * jm, m1 = g.mmin + dm * (jm + 1);
  const double Y = y0 + ((r[0] - sum_ym[kbin]) / (sum_ym[kbin + 1] - sum_ym[kbin])) * (y1 - y0);
  const double M = m0 + r[1] * (m1 - m0);
  // getBinX / getBinY: truncation before the division (include/UpcSampler.h:124-133)
  const int ybin = (int)((double)(int)(g.ny * (Y - g.ymin)) / ((g.ymin + dy * g.ny) - g.ymin));
  int mbin = (int)((double)(int)(g.nm * (M - g.mmin)) / ((g.mmin + dm * g.nm) - g.mmin));
  if (mbin >= g.nm) mbin = g.nm - 1;
  // z = cos(theta) from the z pdf of that M row (src/UpcGenerator.cpp:742-757)
  const double* zt = sum_z;
  if (g.pol) {
    const int yb = ybin >= g.ny ? g.ny - 1 : ybin;
    zt = r[3] < ratio[(long)yb * g.nm + mbin] ? sum_zs : sum_zp;
  }
  zt += (long)mbin * (g.nz + 1);
  const long zb = pdf_find2(zt, g.nz, r[2]);
  if (zb < 0) return;
  const double dz = (g.zmax - g.zmin) / g.nz;
  const double zl = g.zmin + dz * zb, zh = g.zmin + dz * (zb + 1);
  const double cost = zl + ((r[2] - zt[zb]) / (zt[zb + 1] - zt[zb])) * (zh - zl);
  // pair momentum (getPairMomentum)
  double px = 0, py = 0, mt = M;
  if (o.nonzero_gam_pt) {
    const double k1 = M / 2 * exp(Y), k2 = M / 2 * exp(-Y);
    const double a1 = 2 * kPi * r[4], a2 = 2 * kPi * r[5];
    const double pt1 = photon_pt(ptcdf, ptmax, k1, gtot, r[6]);
    const double pt2 = photon_pt(ptcdf, ptmax, k2, gtot, r[7]);
    px = pt1 * cos(a1) + pt2 * cos(a2);
    py = pt1 * sin(a1) + pt2 * sin(a2);
    mt = sqrt(M * M + px * px + py * py);
  }
  const double pz = mt * sinh(Y), E = mt * cosh(Y);
  // pairProduction: back-to-back leptons in the pair frame, boosted to the lab
  const double m2 = E * E - (px * px + py * py + pz * pz);
  const double pmag = sqrt(m2 / 4 - g.mlep * g.mlep);
  const double theta = acos(cost);
  const double phi = 2 * kPi * r[8];
  const double vx = pmag * sin(theta) * cos(phi), vy = pmag * sin(theta) * sin(phi), vz = pmag * cos(theta);
  const double el = sqrt(pmag * pmag + g.mlep * g.mlep);
  double l1[4] = {vx, vy, vz, el}, l2[4] = {-vx, -vy, -vz, el};
  boost(px / E, py / E, pz / E, l1);
  boost(px / E, py / E, pz / E, l2);
  const int sign1 = (2 * r[9] - 1) > 0 ? 1 : -1;
  for (int c = 0; c < 4; ++c) {
    P[c] = l1[c];
    P[4 + c] = l2[c];
  }
  out_pdg[2 * i] = sign1 * g.pdg;
  out_pdg[2 * i + 1] = -sign1 * g.pdg;
  out_pass[i] = (pass_cuts(l1, o) && pass_cuts(l2, o)) ? 1 : 2;
}

static inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

int build_pt_tables(upc_ctx* c, double gtot, hipStream_t s)
{
  PtTables* t = (PtTables*)c->events;
  if (!t) {
    t = new PtTables();
    c->events = t;
  }
  if (!t->cdf) UPC_HIP(hipMalloc(&t->cdf, sizeof(double) * (size_t)kPtEps * (kPtBins + 1)));
  t->gtot = gtot;
  t->ptmax = 6. * kHc / c->cfg.R;
  hipLaunchKernelGGL(k_pt_tables, dim3(kPtEps), dim3(256), 0, s, c->d_ff, t->ptmax, t->cdf);
  UPC_HIP(hipGetLastError());
  return 0;
}

void free_events(upc_ctx* c)
{
  PtTables* t = (PtTables*)c->events;
  if (!t) return;
  if (t->cdf) (void)hipFree(t->cdf);
  delete t;
  c->events = nullptr;
}

int launch_generate_pairs(upc_ctx* c, const double* d_sum_ym, const double* d_sum_z, const double* d_sum_zs,
                          const double* d_sum_zp, const double* d_ratio, const upc_event_opts* opts,
                          unsigned long long seed, unsigned long long offset, long long n, double* d_p, int* d_pdg,
                          int* d_pass, hipStream_t s)
{
  if (n <= 0) return 0;
  const upc_config& cf = c->cfg;
  PtTables* t = (PtTables*)c->events;
  if (opts->nonzero_gam_pt && (!t || !t->cdf)) {
    set_error("photon pt tables not built: call upc_photon_pt_tables first");
    return UPC_ESTATE;
  }
  if (cf.use_polarized ? (!d_sum_zs || !d_sum_zp || !d_ratio) : !d_sum_z) {
    set_error("missing z sampler tables");
    return UPC_EINVAL;
  }
  EvGrid g{cf.nm, cf.ny, cf.nz, cf.use_polarized, cf.proc_id, cf.mmin, cf.mmax, cf.ymin, cf.ymax, cf.zmin, cf.zmax,
           cf.proc_id == 11 ? kMEl : (cf.proc_id == 13 ? kMMu : kMTau)};
  hipLaunchKernelGGL(k_generate_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, *opts, d_sum_ym, d_sum_z,
                     d_sum_zs, d_sum_zp, d_ratio, t ? t->cdf : nullptr, t ? t->ptmax : 0.0, t ? t->gtot : 1.0, seed,
                     offset, n, d_p, d_pdg, d_pass);
  UPC_HIP(hipGetLastError());
  return 0;
}

// photon pt draws alone (for tests of the getPhotonPt distribution)
__global__ void k_photon_pt(const double* __restrict__ cdf, double ptmax, double gtot, const double* __restrict__ E,
                            unsigned long long seed, long long n, double* __restrict__ out)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Philox4x32 ph;
  ph.block(seed, (unsigned long long)i, 7);
  out[i] = photon_pt(cdf, ptmax, E[i], gtot, philox_u01(ph.c[0], ph.c[1]));
}

int launch_photon_pt(upc_ctx* c, const double* d_E, long long n, unsigned long long seed, double* d_out, hipStream_t s)
{
  PtTables* t = (PtTables*)c->events;
  if (!t || !t->cdf) {
    set_error("photon pt tables not built: call upc_photon_pt_tables first");
    return UPC_ESTATE;
  }
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_photon_pt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t->cdf, t->ptmax, t->gtot, d_E,
                     seed, n, d_out);
  UPC_HIP(hipGetLastError());
  return 0;
}

}  // namespace upc
