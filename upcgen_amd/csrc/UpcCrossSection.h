// UpcCrossSection: host facade of the two-photon luminosity / nuclear cross
// section path, with the public interface of upcgen's class of the same name
// (include/UpcCrossSection.h:65-240).  Every computation runs on the GPU
// through the C ABI of include/upc_lumi.h; this class keeps the reference's
// state (class-static beam/nucleus parameters, grid members, flags), its
// call sequence (init -> prepareTwoPhotonLumi -> calcNucCrossSectionYM) and
// its file protocol (lock file + cached luminosity grid in lumiFileDirectory).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/upc_lumi.h"
#include "UpcElemProcess.h"

namespace phys_consts
{
// include/UpcPhysConstants.h:26-32
constexpr double alpha{1.0 / 137.035999074};
constexpr double hc{0.1973269718};
constexpr double mProt{0.9382720813};
constexpr double mNeut{0.939565346};
constexpr double mEl{0.000510998946};
constexpr double mMu{0.1056583745};
constexpr double mTau{1.77686};
}  // namespace phys_consts

// Minimal four-vector with the accessors the reference's uses of
// TLorentzVector need (SetPxPyPzE, Px..E, Pt, M, Eta).
struct UpcLorentzVector {
  double px{0}, py{0}, pz{0}, e{0};
  void SetPxPyPzE(double x, double y, double z, double t)
  {
    px = x;
    py = y;
    pz = z;
    e = t;
  }
  double Px() const { return px; }
  double Py() const { return py; }
  double Pz() const { return pz; }
  double E() const { return e; }
  double Pt() const { return std::sqrt(px * px + py * py); }
  double Mag2() const { return e * e - px * px - py * py - pz * pz; }
  double M() const
  {
    const double m2 = Mag2();
    return m2 < 0 ? -std::sqrt(-m2) : std::sqrt(m2);
  }
  double Eta() const
  {
    const double pt = Pt();
    return pt > 0 ? std::asinh(pz / pt) : (pz >= 0 ? 10e10 : -10e10);
  }
};

class UpcCrossSection
{
 public:
  UpcCrossSection();
  ~UpcCrossSection();
  UpcCrossSection(const UpcCrossSection&) = delete;
  UpcCrossSection& operator=(const UpcCrossSection&) = delete;

  // Woods-Saxon parameters
  inline static double rho0{0.};  // fm^-3
  inline static double R{6.68};   // fm
  inline static double a{0.447};  // fm

  // parameters of the nucleus
  inline static int Z{82};
  inline static int A{208};
  inline static double mNucl{(Z * phys_consts::mProt + (A - Z) * phys_consts::mNeut) / A};

  // beam parameters
  inline static double sqrts{5020.};
  inline static double g1{sqrts / (2. * phys_consts::mProt)};
  inline static double g2{sqrts / (2. * phys_consts::mProt)};

  // 10-point Gauss-Legendre table (include/UpcCrossSection.h:90-111)
  static const int ngi10 = 10;
  double weights10[ngi10]{0.0666713443086881, 0.1494513491505806, 0.2190863625159820, 0.2692667193099963,
                          0.2955242247147529, 0.2955242247147529, 0.2692667193099963, 0.2190863625159820,
                          0.1494513491505806, 0.0666713443086881};
  double abscissas10[ngi10]{-0.9739065285171717, -0.8650633666889845, -0.6794095682990244, -0.4333953941292472,
                            -0.1488743389816312, 0.1488743389816312,  0.4333953941292472,  0.6794095682990244,
                            0.8650633666889845,  0.9739065285171717};

  // photon luminosity calculation parameters
  std::string lumiFileDirectory{"."};
  void setLumiFileDirectory(const std::string& directory) { lumiFileDirectory = directory; }
  const int nb1{120};
  const int nb2{120};

  // cross sections binning
  double zmin{-1};
  double zmax{1};
  int nz{100};
  double mmin{3.56};
  double mmax{50.};
  int nm{1001};
  double ymin{-6.};
  double ymax{6.};
  int ny{121};

  bool doMassCut{false};
  double lowMCut{0};
  double hiMCut{9999};

  // scaling factor
  double factor{Z * Z * phys_consts::alpha / M_PI / M_PI / phys_consts::hc / phys_consts::hc};

  static const int nb{200};

  // Lorentz factor of the photon-pt distribution, frozen at construction
  // from the default beam energy as in the reference
  double gtot;

  // simulation parameters
  bool isPoint{true};
  bool useNonzeroGamPt{true};
  bool usePolarizedCS{false};
  // 1 -- no breakup, 2 -- XNXN, 3 -- 0N0N, 4 -- 0NXN (+ XN0N)
  int breakupMode{1};
  int shadowingOption{0};
  int dghtPDG{};

  inline static int debug{0};
  int numThreads{1};  // kept for the interface: the GPU replaces the OpenMP team

  void setElemProcess(int procID);
  UpcElemProcess* elemProcess{nullptr};

  double alpMass{1.};
  double alpWidth{0.010};

  // ---- additions of this implementation ----
  int lumiForm{UPC_FORM_FP64};   // FP64 quadrature (default), MFMA or surrogate
  int lockTimeoutSec{3600};      // a lock older than this is taken over (reference: waits forever)
  bool strictCacheKey{true};     // recompute when the cached grid was made with other parameters
  upc_ctx* context() const { return ctx_; }
  upc_config config() const;

  void init();

  // methods for cross section calculation (all evaluated on the GPU)
  template <typename ArrayType>
  static double simpson(int n, ArrayType* v, double h);
  double calcWSRho();
  double fluxPoint(double b, double k);
  static double calcFormFac(double Q2);
  double fluxForm(double b, double k);
  double calcTwoPhotonLumi(double M, double Y);
  void calcTwoPhotonLumiPol(double& ns, double& np, double M, double Y);
  void fillCrossSectionZM(std::vector<std::vector<double>>& crossSectionZM, double zmin, double zmax, int nz,
                          double mmin, double mmax, int nm, int flag);
  void calcNucCrossSectionYM(std::vector<std::vector<double>>& crossSectionYM,
                             std::vector<std::vector<double>>& polCSRatio, double& totCS);
  // vector-meson path: photon flux at k = M/2 e^Y weighted by P(b) G_AA(b),
  // and sigma(y) = flux(m, y) sigma_el(y) + flux(m, -y) sigma_el(-y) with the
  // elementary sigma_el(y) of elemProcess (src/UpcCrossSection.cpp:699-748)
  double calcPhotonFlux(double M, double Y);
  void calcNucCrossSectionY(std::vector<std::vector<double>>& crossSectionY,
                            std::vector<std::vector<double>>& csYRatio, double& totCS);
  double calcBreakupProb(double b, int mode);

  // pair momentum with non-zero photon pt (src/UpcCrossSection.cpp:1021-1074):
  // photon pt drawn from the per-energy form-factor distribution on the GPU,
  // azimuths from this object's own generator (the reference's gRandom)
  double getPhotonPt(double ePhot);
  void getPairMomentum(double mPair, double yPair, UpcLorentzVector& pPair);
  void setRandomSeed(unsigned long long seed) { rngState_ = seed; }

  void prepareGAA();
  void prepareBreakupProb();
  void prepareFormFac();
  void prepareTwoPhotonLumi();

  // file names of the lock and of the cached grid
  std::string lockFileName() const;
  std::string lumiFileName() const;

 private:
  upc_ctx* ctx_{nullptr};
  double* dbuf_{nullptr};  // small device buffer for per-point calls
  void ensureContext();
  void evalPoints(int what, const double* a, const double* b, int n, double* out);
  bool readLumiCache(std::vector<double>& l0, std::vector<double>& l1, bool requireKey) const;
  void writeLumiCache(const std::vector<double>& l0, const std::vector<double>& l1) const;
  inline static UpcCrossSection* current_{nullptr};
  bool ptTables_{false};
  unsigned long long ptDraws_{0};     // photon-pt draws so far (Philox stream per draw)
  unsigned long long rngState_{12345};
  double uniform01();                 // splitmix64 stream for the azimuths
};

template <typename ArrayType>
double UpcCrossSection::simpson(int n, ArrayType* v, double h)
{
  // src/UpcCrossSection.cpp:139-150: with even n the closing weights are 2, 1
  double sum = v[0] + v[n - 1];
  for (int i = 1; i < n - 1; i += 2) sum += 4. * v[i];
  for (int i = 2; i < n - 1; i += 2) sum += 2. * v[i];
  return sum * h / 3.;
}
