// UpcCrossSection facade (see UpcCrossSection.h).  Reference behaviour kept:
//   init                  src/UpcCrossSection.cpp:116-137
//   prepareTwoPhotonLumi  src/UpcCrossSection.cpp:463-592 (lock, cache probe, grid, unlock)
//   calcNucCrossSectionYM src/UpcCrossSection.cpp:594-698 (re-reads the cached grid)
#include "UpcCrossSection.h"

#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <stdexcept>

#include "UpcLog.h"

namespace
{
void hipCheck(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    UPC_LOG_FATAL << what << ": " << hipGetErrorString(e);
    std::_Exit(-1);
  }
}

void abiCheck(int rc, const char* what)
{
  if (rc != 0) {
    UPC_LOG_FATAL << what << " failed: " << upc_last_error();
    std::_Exit(-1);
  }
}

// Cached luminosity grid: fixed header (the parameters the grid depends on)
// followed by nm*ny doubles per grid, [im][iy] = TH2D bin (im + 1, iy + 1) of
// the reference's twoPhotonLumi[Pol].root.
struct LumiCacheHeader {
  char magic[8];
  int32_t version, polarized, Z, A, flux_point, breakup_mode, nm, ny;
  double R, a, sqrts, mmin, mmax, ymin, ymax;
  int32_t form, pad;  // luminosity form that made the grid (UPC_FORM_*)
};

LumiCacheHeader makeHeader(const upc_config& c, int form)
{
  LumiCacheHeader h{};
  std::memcpy(h.magic, "UPCLUMI2", 8);
  h.version = 2;
  h.form = form;
  h.polarized = c.use_polarized;
  h.Z = c.Z;
  h.A = c.A;
  h.flux_point = c.flux_point;
  h.breakup_mode = c.breakup_mode;
  h.nm = c.nm;
  h.ny = c.ny;
  h.R = c.R;
  h.a = c.a;
  h.sqrts = c.sqrts;
  h.mmin = c.mmin;
  h.mmax = c.mmax;
  h.ymin = c.ymin;
  h.ymax = c.ymax;
  return h;
}

bool sameKey(const LumiCacheHeader& x, const LumiCacheHeader& y)
{
  return x.polarized == y.polarized && x.Z == y.Z && x.A == y.A && x.flux_point == y.flux_point &&
         x.breakup_mode == y.breakup_mode && x.nm == y.nm && x.ny == y.ny && x.R == y.R && x.a == y.a &&
         x.sqrts == y.sqrts && x.mmin == y.mmin && x.mmax == y.mmax && x.ymin == y.ymin && x.ymax == y.ymax &&
         x.form == y.form;
}
}  // namespace

// ---------------------------------------------------------------------------
UpcTwoPhotonDilep::UpcTwoPhotonDilep(int pdg)
{
  // src/UpcTwoPhotonDilep.cpp:29-44
  partPDG = pdg;
  isCharged = true;
  if (pdg == 11) mPart = phys_consts::mEl;
  if (pdg == 13) mPart = phys_consts::mMu;
  if (pdg == 15) mPart = phys_consts::mTau;
}

UpcTwoPhotonDilep::~UpcTwoPhotonDilep()
{
  if (dbuf_) (void)hipFree(dbuf_);
}

double UpcTwoPhotonDilep::eval(int flag, double m, double z)
{
  if (!ctx_) {
    UPC_LOG_FATAL << "UpcTwoPhotonDilep used before UpcCrossSection::init()";
    std::_Exit(-1);
  }
  if (!dbuf_) hipCheck(hipMalloc(&dbuf_, 4 * sizeof(double)), "hipMalloc");
  const double h[2] = {m, z};
  hipCheck(hipMemcpy(dbuf_, h, sizeof(h), hipMemcpyHostToDevice), "hipMemcpy");
  abiCheck(upc_elem_cs(ctx_, flag, dbuf_, dbuf_ + 1, 1, dbuf_ + 2, nullptr), "upc_elem_cs");
  double out = 0;
  hipCheck(hipMemcpy(&out, dbuf_ + 2, sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
  return out;
}

double UpcTwoPhotonDilep::calcCrossSectionM(double m) { return eval(0, m, 0.); }
double UpcTwoPhotonDilep::calcCrossSectionMPolS(double m) { return eval(1, m, 0.); }
double UpcTwoPhotonDilep::calcCrossSectionMPolPS(double m) { return eval(2, m, 0.); }
double UpcTwoPhotonDilep::calcCrossSectionZM(double z, double m) { return eval(3, m, z); }
double UpcTwoPhotonDilep::calcCrossSectionZMPolS(double z, double m) { return eval(4, m, z); }
double UpcTwoPhotonDilep::calcCrossSectionZMPolPS(double z, double m) { return eval(5, m, z); }

// ---------------------------------------------------------------------------
UpcCrossSection::UpcCrossSection()
{
  // src/UpcCrossSection.cpp:51-54
  gtot = std::cosh((std::acosh(g1) + std::acosh(g2)) / 2.);
  current_ = this;
}

UpcCrossSection::~UpcCrossSection()
{
  delete elemProcess;
  if (dbuf_) (void)hipFree(dbuf_);
  if (ctx_) upc_ctx_destroy(ctx_);
  if (current_ == this) current_ = nullptr;
}

void UpcCrossSection::setElemProcess(int procID)
{
  // src/UpcCrossSection.cpp:67-114; only the dilepton processes are part of
  // the luminosity engine
  switch (procID) {
    case 11:
    case 13:
    case 15:
      delete elemProcess;
      elemProcess = new UpcTwoPhotonDilep(procID);
      break;
    default:
      UPC_LOG_FATAL << "Process ID " << procID
                    << " is not supported by this engine (dileptons: 11, 13, 15)! Exiting...";
      std::_Exit(-1);
  }
}

upc_config UpcCrossSection::config() const
{
  upc_config c{};
  c.Z = Z;
  c.A = A;
  c.R = R;
  c.a = a;
  c.sqrts = sqrts;
  c.flux_point = isPoint ? 1 : 0;
  c.breakup_mode = breakupMode;
  c.use_polarized = usePolarizedCS ? 1 : 0;
  c.nm = nm;
  c.mmin = mmin;
  c.mmax = mmax;
  c.ny = ny;
  c.ymin = ymin;
  c.ymax = ymax;
  c.nz = nz;
  c.zmin = zmin;
  c.zmax = zmax;
  c.proc_id = elemProcess ? elemProcess->partPDG : 13;
  auto* dl = dynamic_cast<UpcTwoPhotonDilep*>(elemProcess);
  c.lep_a = dl ? dl->aLep : 0.0;
  return c;
}

void UpcCrossSection::ensureContext()
{
  if (ctx_) return;
  if (!elemProcess) setElemProcess(13);
  const upc_config cfg = config();
  abiCheck(upc_ctx_create(&cfg, nullptr, &ctx_), "upc_ctx_create");
  hipCheck(hipMalloc(&dbuf_, 8 * sizeof(double)), "hipMalloc");
  upc_info inf{};
  abiCheck(upc_ctx_info(ctx_, &inf), "upc_ctx_info");
  rho0 = inf.rho0;
  if (auto* dl = dynamic_cast<UpcTwoPhotonDilep*>(elemProcess)) dl->attach(ctx_);
  current_ = this;
}

void UpcCrossSection::init()
{
  UPC_LOG_INFO << "Initializing caches ...";
  factor = Z * Z * phys_consts::alpha / M_PI / M_PI / phys_consts::hc / phys_consts::hc;
  mNucl = (Z * phys_consts::mProt + (A - Z) * phys_consts::mNeut) / A;
  if (ctx_) {
    upc_ctx_destroy(ctx_);
    ctx_ = nullptr;
    (void)hipFree(dbuf_);
    dbuf_ = nullptr;
  }
  // calcWSRho, prepareGAA, prepareFormFac, prepareBreakupProb: one GPU pass
  ensureContext();
  prepareTwoPhotonLumi();
}

double UpcCrossSection::calcWSRho()
{
  ensureContext();
  return rho0;
}

void UpcCrossSection::prepareGAA() { ensureContext(); }
void UpcCrossSection::prepareFormFac() { ensureContext(); }
void UpcCrossSection::prepareBreakupProb()
{
  ensureContext();
  abiCheck(upc_ctx_rebuild_tables(ctx_, nullptr), "upc_ctx_rebuild_tables");
}

// what: 0 flux point, 1 flux form, 2 form factor, 3 breakup mode (int in b[0]), 4 lumi, 5 lumi pol
void UpcCrossSection::evalPoints(int what, const double* x, const double* y, int n, double* out)
{
  ensureContext();
  double h[2] = {x[0], y ? y[0] : 0.0};
  hipCheck(hipMemcpy(dbuf_, h, sizeof(h), hipMemcpyHostToDevice), "hipMemcpy");
  double* o = dbuf_ + 2;
  switch (what) {
    case 0: abiCheck(upc_flux_point(ctx_, dbuf_, dbuf_ + 1, n, o, nullptr), "upc_flux_point"); break;
    case 1: abiCheck(upc_flux_form(ctx_, dbuf_, dbuf_ + 1, n, o, nullptr, nullptr), "upc_flux_form"); break;
    case 2: abiCheck(upc_form_factor(ctx_, dbuf_, n, 0, o, nullptr), "upc_form_factor"); break;
    case 3:
      abiCheck(upc_breakup_prob(ctx_, dbuf_, n, (int)y[0], 0, o, nullptr), "upc_breakup_prob");
      break;
    case 4: abiCheck(upc_lumi_points(ctx_, dbuf_, dbuf_ + 1, n, o, nullptr, nullptr, nullptr), "upc_lumi_points"); break;
    default:
      abiCheck(upc_lumi_points(ctx_, dbuf_, dbuf_ + 1, n, nullptr, o, o + 1, nullptr), "upc_lumi_points");
      break;
  }
  hipCheck(hipMemcpy(out, o, (what == 5 ? 2 : 1) * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
}

double UpcCrossSection::fluxPoint(double b, double k)
{
  double o;
  evalPoints(0, &b, &k, 1, &o);
  return o;
}

double UpcCrossSection::fluxForm(double b, double k)
{
  double o;
  evalPoints(1, &b, &k, 1, &o);
  return o;
}

double UpcCrossSection::calcFormFac(double Q2)
{
  if (!current_) {
    UPC_LOG_FATAL << "calcFormFac needs an UpcCrossSection instance";
    std::_Exit(-1);
  }
  double o;
  current_->evalPoints(2, &Q2, nullptr, 1, &o);
  return o;
}

double UpcCrossSection::calcBreakupProb(double b, int mode)
{
  double o;
  const double md = mode;
  evalPoints(3, &b, &md, 1, &o);
  return o;
}

double UpcCrossSection::calcTwoPhotonLumi(double M, double Y)
{
  if (usePolarizedCS) {
    // the context computes the polarised pair; their sum is not the
    // unpolarised luminosity (sign convention of src/UpcCrossSection.cpp:257 vs :317)
    UPC_LOG_FATAL << "calcTwoPhotonLumi on a polarised configuration: use calcTwoPhotonLumiPol";
    std::_Exit(-1);
  }
  double o;
  evalPoints(4, &M, &Y, 1, &o);
  return o;
}

void UpcCrossSection::calcTwoPhotonLumiPol(double& ns, double& np, double M, double Y)
{
  if (!usePolarizedCS) {
    UPC_LOG_FATAL << "calcTwoPhotonLumiPol on an unpolarised configuration";
    std::_Exit(-1);
  }
  double o[2];
  evalPoints(5, &M, &Y, 1, o);
  ns = o[0];
  np = o[1];
}

void UpcCrossSection::fillCrossSectionZM(std::vector<std::vector<double>>& crossSectionZM, double zmin_,
                                         double zmax_, int nz_, double mmin_, double mmax_, int nm_, int flag)
{
  ensureContext();
  double* d = nullptr;
  hipCheck(hipMalloc(&d, sizeof(double) * (size_t)nm_ * nz_), "hipMalloc");
  abiCheck(upc_cs_zm(ctx_, zmin_, zmax_, nz_, mmin_, mmax_, nm_, flag, d, nullptr), "upc_cs_zm");
  std::vector<double> h((size_t)nm_ * nz_);
  hipCheck(hipMemcpy(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
  (void)hipFree(d);
  crossSectionZM.resize(nm_);
  for (int im = 0; im < nm_; ++im) crossSectionZM[im].assign(h.begin() + (size_t)im * nz_, h.begin() + (size_t)(im + 1) * nz_);
}

std::string UpcCrossSection::lockFileName() const
{
  return lumiFileDirectory + "/.lumiIsCalculated" + (usePolarizedCS ? "Pol" : "");
}

std::string UpcCrossSection::lumiFileName() const
{
  return lumiFileDirectory + "/twoPhotonLumi" + (usePolarizedCS ? "Pol" : "") + ".upcbin";
}

bool UpcCrossSection::readLumiCache(std::vector<double>& l0, std::vector<double>& l1, bool requireKey) const
{
  std::ifstream f(lumiFileName(), std::ios::binary);
  if (!f) return false;
  LumiCacheHeader h{};
  f.read(reinterpret_cast<char*>(&h), sizeof(h));
  if (!f || std::memcmp(h.magic, "UPCLUMI2", 8) != 0) {
    UPC_LOG_WARNING << "Cached luminosity file " << lumiFileName() << " is not readable: recomputing";
    return false;
  }
  const LumiCacheHeader want = makeHeader(config(), lumiForm);
  if (!sameKey(h, want)) {
    if (requireKey) {
      UPC_LOG_WARNING << "Cached luminosity " << lumiFileName()
                      << " was made with different parameters: recomputing";
      return false;
    }
    if (h.nm != nm || h.ny != ny || h.polarized != (usePolarizedCS ? 1 : 0)) return false;
    UPC_LOG_WARNING << "Reusing cached luminosity made with different parameters (as the reference does)";
  }
  const size_t n = (size_t)nm * ny;
  l0.resize(n);
  f.read(reinterpret_cast<char*>(l0.data()), sizeof(double) * n);
  if (usePolarizedCS) {
    l1.resize(n);
    f.read(reinterpret_cast<char*>(l1.data()), sizeof(double) * n);
  }
  return (bool)f;
}

void UpcCrossSection::writeLumiCache(const std::vector<double>& l0, const std::vector<double>& l1) const
{
  const std::string tmp = lumiFileName() + ".tmp";
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    const LumiCacheHeader h = makeHeader(config(), lumiForm);
    f.write(reinterpret_cast<const char*>(&h), sizeof(h));
    f.write(reinterpret_cast<const char*>(l0.data()), sizeof(double) * l0.size());
    if (usePolarizedCS) f.write(reinterpret_cast<const char*>(l1.data()), sizeof(double) * l1.size());
    if (!f) {
      UPC_LOG_WARNING << "Could not write " << tmp;
      return;
    }
  }
  std::rename(tmp.c_str(), lumiFileName().c_str());
}

void UpcCrossSection::prepareTwoPhotonLumi()
{
  ensureContext();
  // wait here if another generator is working on this (src/UpcCrossSection.cpp:465-478),
  // taking over a lock older than lockTimeoutSec
  const std::string fnlock = lockFileName();
  while (true) {
    const int fd = open(fnlock.c_str(), O_CREAT | O_EXCL | O_WRONLY, 0644);
    if (fd >= 0) {
      close(fd);
      break;
    }
    struct stat st {};
    if (lockTimeoutSec >= 0 && stat(fnlock.c_str(), &st) == 0 &&
        std::difftime(std::time(nullptr), st.st_mtime) > lockTimeoutSec) {
      UPC_LOG_WARNING << "Lock file " << fnlock << " is older than " << lockTimeoutSec << " s: taking it over";
      std::remove(fnlock.c_str());
      continue;
    }
    sleep(1);
  }

  std::vector<double> l0, l1;
  if (readLumiCache(l0, l1, strictCacheKey)) {
    UPC_LOG_INFO << "Found pre-calculated " << (usePolarizedCS ? "polarized" : "unpolarized") << " 2D luminosity";
  } else {
    UPC_LOG_INFO << "Precalculated 2D luminosity is not found. Starting all over...";
    const size_t n = (size_t)nm * ny;
    double* d = nullptr;
    hipCheck(hipMalloc(&d, sizeof(double) * n * 2), "hipMalloc");
    if (usePolarizedCS)
      abiCheck(upc_lumi_grid(ctx_, lumiForm, nullptr, d, d + n, nullptr), "upc_lumi_grid");
    else
      abiCheck(upc_lumi_grid(ctx_, lumiForm, d, nullptr, nullptr, nullptr), "upc_lumi_grid");
    l0.resize(n);
    hipCheck(hipMemcpy(l0.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    if (usePolarizedCS) {
      l1.resize(n);
      hipCheck(hipMemcpy(l1.data(), d + n, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    }
    (void)hipFree(d);
    upc_info inf{};
    upc_ctx_info(ctx_, &inf);
    if (inf.qags_bad)
      UPC_LOG_WARNING << inf.qags_bad << " form-factor flux integrals ended with a QAGS error"
                      << " (the reference aborts there)";
    writeLumiCache(l0, l1);
    UPC_LOG_INFO << "Two-photon luminosity was written to " << lumiFileName();
  }
  if (std::remove(fnlock.c_str()) != 0) UPC_LOG_WARNING << "Lock file " << fnlock << " was not properly removed!";
}

void UpcCrossSection::calcNucCrossSectionYM(std::vector<std::vector<double>>& crossSectionYM,
                                            std::vector<std::vector<double>>& polCSRatio, double& totCS)
{
  UPC_LOG_INFO << "Calculating nuclear cross section...";
  ensureContext();
  std::vector<double> l0, l1;
  if (!readLumiCache(l0, l1, false)) {
    UPC_LOG_FATAL << "Two-photon luminosity file " << lumiFileName() << " is missing";
    std::_Exit(-1);
  }
  const size_t n = (size_t)nm * ny;
  double* d = nullptr;
  hipCheck(hipMalloc(&d, sizeof(double) * n * 4), "hipMalloc");
  double *dL = d, *dLp = d + n, *dcs = d + 2 * n, *dr = d + 3 * n;
  hipCheck(hipMemcpy(dL, l0.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
  if (usePolarizedCS) {
    hipCheck(hipMemcpy(dLp, l1.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    abiCheck(upc_nuc_cs_ym(ctx_, nullptr, dL, dLp, dcs, dr, &totCS, nullptr), "upc_nuc_cs_ym");
  } else {
    abiCheck(upc_nuc_cs_ym(ctx_, dL, nullptr, nullptr, dcs, nullptr, &totCS, nullptr), "upc_nuc_cs_ym");
  }
  std::vector<double> h(n), hr;
  hipCheck(hipMemcpy(h.data(), dcs, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
  if (usePolarizedCS) {
    hr.resize(n);
    hipCheck(hipMemcpy(hr.data(), dr, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
  }
  (void)hipFree(d);
  crossSectionYM.assign(ny, std::vector<double>(nm));
  if (usePolarizedCS) polCSRatio.assign(ny, std::vector<double>(nm));
  for (int iy = 0; iy < ny; ++iy)
    for (int im = 0; im < nm; ++im) {
      crossSectionYM[iy][im] = h[(size_t)iy * nm + im];
      if (usePolarizedCS) polCSRatio[iy][im] = hr[(size_t)iy * nm + im];
    }
  UPC_LOG_INFO << "Calculating nuclear cross section...Done!";
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.6f", totCS);
  UPC_LOG_INFO << "Total nuclear cross section = " << buf << " mb";
}
