// UpcCrossSection facade (see UpcCrossSection.h).  Reference behaviour kept:
//   init                  src/UpcCrossSection.cpp:116-137
//   prepareTwoPhotonLumi  src/UpcCrossSection.cpp:463-592 (lock, cache probe, grid, unlock)
//   calcNucCrossSectionYM src/UpcCrossSection.cpp:594-698 (re-reads the cached grid)
#include "UpcCrossSection.h"

#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/file.h>
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
// The luminosity lock this process holds while it computes the grid: the
// lock file (created, or taken over when stale) and an exclusive flock on it.
// Fatal exits release it first, so that later runs do not wait for it.
struct HeldLock {
  std::string path;
  int fd = -1;
};
HeldLock g_lock;

void releaseHeldLock()
{
  if (g_lock.fd < 0) return;
  if (std::remove(g_lock.path.c_str()) != 0)
    UPC_LOG_WARNING << "Lock file " << g_lock.path << " was not properly removed!";
  close(g_lock.fd);  // drops the flock
  g_lock = HeldLock{};
}

void hipCheck(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    UPC_LOG_FATAL << what << ": " << hipGetErrorString(e);
    releaseHeldLock();
    std::_Exit(-1);
  }
}

void abiCheck(int rc, const char* what)
{
  if (rc != 0) {
    UPC_LOG_FATAL << what << " failed: " << upc_last_error();
    releaseHeldLock();
    std::_Exit(-1);
  }
}

bool sameFile(int fd, const std::string& path)
{
  struct stat a {}, b {};
  return fstat(fd, &a) == 0 && stat(path.c_str(), &b) == 0 && a.st_dev == b.st_dev && a.st_ino == b.st_ino;
}

// Waits for and takes the luminosity lock at `path` (src/UpcCrossSection.cpp:
// 465-478: the lock file's existence means another generator is computing).
// The holder keeps an exclusive flock on the file, so a lock is stale only
// when nobody holds its flock (the holder died, or it was left by a process
// that does not take one) and it is older than timeout_s (< 0: never).  A
// stale lock is taken over by acquiring its flock and refreshing its time, so
// of several waiters exactly one takes it; a holder that ends removes the file
// before dropping the flock, and a waiter that got the flock of a file no
// longer at `path` starts over.
void acquireLock(const std::string& path, int timeout_s)
{
  while (true) {
    int fd = open(path.c_str(), O_CREAT | O_EXCL | O_RDWR, 0644);
    if (fd >= 0) {
      if (flock(fd, LOCK_EX) == 0 && sameFile(fd, path)) {
        g_lock = HeldLock{path, fd};
        return;
      }
      close(fd);
      continue;
    }
    fd = open(path.c_str(), O_RDWR);
    if (fd < 0) continue;  // removed meanwhile: try to create it again
    struct stat st {};
    const bool old = timeout_s >= 0 && fstat(fd, &st) == 0 &&
                     std::difftime(std::time(nullptr), st.st_mtime) > timeout_s;
    if (old && flock(fd, LOCK_EX | LOCK_NB) == 0) {
      if (sameFile(fd, path)) {
        futimens(fd, nullptr);
        UPC_LOG_WARNING << "Lock file " << path << " is older than " << timeout_s
                        << " s and nobody holds it: taking it over";
        g_lock = HeldLock{path, fd};
        return;
      }
    }
    close(fd);
    sleep(1);
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
  // the context's own elementary tables cover the dileptons; another process
  // (a caller's vector meson for calcNucCrossSectionY) leaves them at the muon's
  const int pdg = elemProcess ? elemProcess->partPDG : 13;
  c.proc_id = (pdg == 11 || pdg == 13 || pdg == 15) ? pdg : 13;
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
    case 6: abiCheck(upc_photon_flux(ctx_, dbuf_, dbuf_ + 1, n, o, nullptr), "upc_photon_flux"); break;
    case 7: abiCheck(upc_photon_pt(ctx_, dbuf_, n, (unsigned long long)y[0], o, nullptr), "upc_photon_pt"); break;
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

double UpcCrossSection::calcPhotonFlux(double M, double Y)
{
  double o;
  evalPoints(6, &M, &Y, 1, &o);
  return o;
}

void UpcCrossSection::calcNucCrossSectionY(std::vector<std::vector<double>>& crossSectionY,
                                           std::vector<std::vector<double>>& csYRatio, double& totCS)
{
  UPC_LOG_INFO << "Calculating nuclear cross section...";
  ensureContext();
  const double dy = (ymax - ymin) / ny;
  std::vector<double> h(4 * (size_t)ny);  // sigma(y), sigma(-y), cs, ratio
  for (int iy = 0; iy < ny; ++iy) {
    const double y = ymin + dy * iy;
    h[iy] = elemProcess->calcCrossSectionY(y);
    h[ny + iy] = elemProcess->calcCrossSectionY(-y);
  }
  double* d = nullptr;
  hipCheck(hipMalloc(&d, sizeof(double) * h.size()), "hipMalloc");
  hipCheck(hipMemcpy(d, h.data(), sizeof(double) * 2 * ny, hipMemcpyHostToDevice), "hipMemcpy");
  abiCheck(upc_nuc_cs_y(ctx_, elemProcess->mPart, d, d + ny, d + 2 * ny, d + 3 * ny, &totCS, nullptr),
           "upc_nuc_cs_y");
  hipCheck(hipMemcpy(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
  (void)hipFree(d);
  crossSectionY.resize(ny);
  csYRatio.resize(ny);
  for (int iy = 0; iy < ny; ++iy) {
    crossSectionY[iy].resize(1);
    csYRatio[iy].resize(1);
    crossSectionY[iy][0] = h[2 * ny + iy];
    csYRatio[iy][0] = h[3 * ny + iy];
  }
}

double UpcCrossSection::uniform01()
{
  // splitmix64 (Steele, Lea, Flood 2014): 53 random bits per draw
  unsigned long long z = (rngState_ += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}

double UpcCrossSection::getPhotonPt(double ePhot)
{
  ensureContext();
  if (!ptTables_) {
    abiCheck(upc_photon_pt_tables(ctx_, gtot, nullptr), "upc_photon_pt_tables");
    ptTables_ = true;
  }
  const double seed = (double)(++ptDraws_);  // one Philox stream per draw
  double o;
  evalPoints(7, &ePhot, &seed, 1, &o);
  return o;
}

void UpcCrossSection::getPairMomentum(double mPair, double yPair, UpcLorentzVector& pPair)
{
  if (!useNonzeroGamPt) {
    const double mtPair = mPair;  // pair pt = 0
    pPair.SetPxPyPzE(0., 0., mtPair * std::sinh(yPair), mtPair * std::cosh(yPair));
    return;
  }
  const double k1 = mPair / 2 * std::exp(yPair);
  const double k2 = mPair / 2 * std::exp(-yPair);
  const double angle1 = 2 * M_PI * uniform01();
  const double angle2 = 2 * M_PI * uniform01();
  const double pt1 = getPhotonPt(k1);
  const double pt2 = getPhotonPt(k2);
  const double px = pt1 * std::cos(angle1) + pt2 * std::cos(angle2);
  const double py = pt1 * std::sin(angle1) + pt2 * std::sin(angle2);
  const double pt = std::sqrt(px * px + py * py);
  const double mtPair = std::sqrt(mPair * mPair + pt * pt);
  pPair.SetPxPyPzE(px, py, mtPair * std::sinh(yPair), mtPair * std::cosh(yPair));
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
  // a file of this process's own, renamed into place when complete
  const std::string tmp = lumiFileName() + ".tmp." + std::to_string((long)getpid());
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    const LumiCacheHeader h = makeHeader(config(), lumiForm);
    f.write(reinterpret_cast<const char*>(&h), sizeof(h));
    f.write(reinterpret_cast<const char*>(l0.data()), sizeof(double) * l0.size());
    if (usePolarizedCS) f.write(reinterpret_cast<const char*>(l1.data()), sizeof(double) * l1.size());
    if (!f) {
      UPC_LOG_WARNING << "Could not write " << tmp;
      std::remove(tmp.c_str());
      return;
    }
  }
  if (std::rename(tmp.c_str(), lumiFileName().c_str()) != 0) {
    UPC_LOG_WARNING << "Could not rename " << tmp << " to " << lumiFileName();
    std::remove(tmp.c_str());
  }
}

void UpcCrossSection::prepareTwoPhotonLumi()
{
  ensureContext();
  // wait here if another generator is working on this (src/UpcCrossSection.cpp:465-478),
  // taking over a stale lock (see acquireLock)
  acquireLock(lockFileName(), lockTimeoutSec);

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
    if (inf.qags_bad) {
      // The reference process stops at the first of these: GSL's default
      // error handler aborts on a QAGS error (GSL_EROUND, ...) and on a
      // form-factor spline evaluation past the table's last knot (GSL_EDOM),
      // whose NaN then fills the bins that use the flux row.  Here those
      // bins are emptied so that the cross sections and samplers stay usable.
      size_t nnan = 0;
      for (std::vector<double>* v : {&l0, &l1})
        for (double& x : *v)
          if (std::isnan(x)) {
            x = 0.;
            ++nnan;
          }
      UPC_LOG_ERROR << inf.qags_bad << " form-factor flux integrals ended with a QAGS error or a form-factor"
                    << " spline domain error (GSL would abort the reference here); " << nnan
                    << " luminosity bins with NaN are set to 0";
    }
    writeLumiCache(l0, l1);
    UPC_LOG_INFO << "Two-photon luminosity was written to " << lumiFileName();
  }
  releaseHeldLock();
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
