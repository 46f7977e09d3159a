// UpcGenerator over the GPU engine.  Reference behaviour kept:
//   setParameterValue        src/UpcGenerator.cpp:183-305 (same keys)
//   configGeneratorFromFile  src/UpcGenerator.cpp:318-346 (same parsing)
//   init                     src/UpcGenerator.cpp:57-181 (dilepton checks, mmin clamp)
//   computeNuclXsection      src/UpcGenerator.cpp:630-713
//   generateEvent(s)         src/UpcGenerator.cpp:715-926 (dileptons, no Pythia)
#include "UpcGenerator.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <ctime>
#include <sstream>
#include <sys/stat.h>

#include "UpcLog.h"

int UpcGenerator::debug = 0;

namespace
{
void hipOk(hipError_t e, const char* what)
{
  if (e != hipSuccess) {
    UPC_LOG_FATAL << what << ": " << hipGetErrorString(e);
    std::_Exit(-1);
  }
}

void abiOk(int rc, const char* what)
{
  if (rc != 0) {
    UPC_LOG_FATAL << what << " failed: " << upc_last_error();
    std::_Exit(-1);
  }
}

bool fileExists(const std::string& f)
{
  struct stat st {};
  return stat(f.c_str(), &st) == 0;
}
}  // namespace

UpcGenerator::UpcGenerator()
{
  nucProcessCS = new UpcCrossSection();
  totCS = 0.0;
  fidCS = 0.0;
  setCollisionSystem(5020., 82, 208);
}

UpcGenerator::~UpcGenerator()
{
  delete samplerCsYM;
  for (auto* s : samplersCsZ) delete s;
  for (auto* s : samplersCsSZ) delete s;
  for (auto* s : samplersCsPsZ) delete s;
  freeDevice();
  delete writerHepMC;
  delete nucProcessCS;
}

void UpcGenerator::freeDevice()
{
  for (double** p : {&d_sum_ym, &d_sum_z, &d_sum_zs, &d_sum_zp, &d_ratio}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

void UpcGenerator::setCollisionSystem(float sqrts, int nucl_z, int nucl_a)
{
  UpcCrossSection::sqrts = sqrts;
  UpcCrossSection::g1 = UpcCrossSection::sqrts / (2. * phys_consts::mProt);
  UpcCrossSection::g2 = UpcCrossSection::sqrts / (2. * phys_consts::mProt);
  UpcCrossSection::Z = nucl_z;
  UpcCrossSection::A = nucl_a;
  UpcCrossSection::mNucl = (nucl_z * phys_consts::mProt + (nucl_a - nucl_z) * phys_consts::mNeut) / nucl_a;
}

void UpcGenerator::setParameterValue(const std::string& parameter, const std::string& parValue)
{
  UpcCrossSection* cs = nucProcessCS;
  if (parameter == "NEVENTS") nEvents = std::stol(parValue);
  if (parameter == "SQRTS") {
    UpcCrossSection::sqrts = std::stod(parValue);
    UpcCrossSection::g1 = UpcCrossSection::sqrts / (2. * phys_consts::mProt);
    UpcCrossSection::g2 = UpcCrossSection::sqrts / (2. * phys_consts::mProt);
  }
  if (parameter == "PROC_ID") procID = std::stoi(parValue);
  if (parameter == "LEP_A") aLep = std::stod(parValue);
  if (parameter == "ALP_MASS") cs->alpMass = std::stod(parValue);
  if (parameter == "ALP_WIDTH") cs->alpWidth = std::stod(parValue);
  if (parameter == "DO_PT_CUT") doPtCut = std::stoi(parValue);
  if (parameter == "PT_MIN") minPt = std::stod(parValue);
  if (parameter == "DO_ETA_CUT") doEtaCut = std::stoi(parValue);
  if (parameter == "ETA_MIN") minEta = std::stod(parValue);
  if (parameter == "ETA_MAX") maxEta = std::stod(parValue);
  if (parameter == "ZMIN") cs->zmin = std::stod(parValue);
  if (parameter == "ZMAX") cs->zmax = std::stod(parValue);
  if (parameter == "MMIN") cs->mmin = std::stod(parValue);
  if (parameter == "MMAX") cs->mmax = std::stod(parValue);
  if (parameter == "YMIN") cs->ymin = std::stod(parValue);
  if (parameter == "YMAX") cs->ymax = std::stod(parValue);
  if (parameter == "BINS_Z") cs->nz = std::stoi(parValue);
  if (parameter == "BINS_M") cs->nm = std::stoi(parValue);
  if (parameter == "BINS_Y") cs->ny = std::stoi(parValue);
  if (parameter == "WS_R") UpcCrossSection::R = std::stod(parValue);
  if (parameter == "WS_A") UpcCrossSection::a = std::stod(parValue);
  if (parameter == "NUCLEUS_Z") UpcCrossSection::Z = std::stoi(parValue);
  if (parameter == "NUCLEUS_A") UpcCrossSection::A = std::stoi(parValue);
  if (parameter == "FLUX_POINT") cs->isPoint = std::stoi(parValue);
  if (parameter == "BREAKUP_MODE") cs->breakupMode = std::stoi(parValue);
  if (parameter == "PYTHIA_VERSION") pythiaVersion = std::stoi(parValue);
  if (parameter == "PYTHIA8_FSR") doFSR = std::stoi(parValue);
  if (parameter == "PYTHIA8_DECAYS") doDecays = std::stoi(parValue);
  if (parameter == "NON_ZERO_GAM_PT") cs->useNonzeroGamPt = std::stoi(parValue);
  if (parameter == "USE_POLARIZED_CS") {
    cs->usePolarizedCS = std::stoi(parValue);
    usePolarizedCS = std::stoi(parValue);
  }
  if (parameter == "SEED") seed = std::stol(parValue);
  if (parameter == "USE_ROOT_OUTPUT") useROOTOut = std::stoi(parValue);
  if (parameter == "USE_HEPMC_OUTPUT") useHepMCOut = std::stoi(parValue);
  if (parameter == "DO_M_CUT") cs->doMassCut = std::stoi(parValue);
  if (parameter == "LOW_M_CUT") cs->lowMCut = std::stod(parValue);
  if (parameter == "HIGH_M_CUT") cs->hiMCut = std::stod(parValue);
  if (parameter == "SHADOWING") cs->shadowingOption = std::stoi(parValue);
  if (parameter == "DECAY_PDG") cs->dghtPDG = std::stoi(parValue);
}

void UpcGenerator::configGeneratorFromFile()
{
  if (fileExists(parFileName)) {
    UPC_LOG_INFO << "Reading parameters from " << parFileName << " ...";
    std::ifstream fInputs(parFileName);
    std::string line, parameter, parValue;
    while (std::getline(fInputs, line)) {
      // the stream takes the line before comments are trimmed; a line that
      // yields no tokens leaves the previous key/value pair to be applied again
      std::istringstream iss(line);
      if (line[0] == '#') continue;
      line.erase(std::find(line.begin(), line.end(), '#'), line.end());
      iss >> parameter >> parValue;
      setParameterValue(parameter, parValue);
    }
    if (!useROOTOut && !useHepMCOut)
      UPC_LOG_WARNING << "Output format not set! Choose ROOT or/and HepMC via flags USE_ROOT_OUTPUT and "
                         "USE_HEPMC_OUTPUT!";
  } else {
    UPC_LOG_WARNING << "Input file not found! Using default parameters...";
  }
}

void UpcGenerator::printParameters()
{
  UpcCrossSection* cs = nucProcessCS;
  UPC_LOG_INFO << "OMP_NTHREADS " << numThreads;
  UPC_LOG_INFO << "NUCLEUS_Z " << UpcCrossSection::Z;
  UPC_LOG_INFO << "NUCLEUS_A " << UpcCrossSection::A;
  UPC_LOG_INFO << "WS_R " << UpcCrossSection::R;
  UPC_LOG_INFO << "WS_A " << UpcCrossSection::a;
  UPC_LOG_INFO << "SQRTS " << UpcCrossSection::sqrts;
  UPC_LOG_INFO << "PROC_ID " << procID;
  UPC_LOG_INFO << "LEP_A " << aLep;
  UPC_LOG_INFO << "NEVENTS " << nEvents;
  UPC_LOG_INFO << "DO_PT_CUT " << doPtCut;
  UPC_LOG_INFO << "PT_MIN " << minPt;
  UPC_LOG_INFO << "DO_ETA_CUT " << doEtaCut;
  UPC_LOG_INFO << "ETA_MIN " << minEta;
  UPC_LOG_INFO << "ETA_MAX " << maxEta;
  UPC_LOG_INFO << "ZMIN " << cs->zmin;
  UPC_LOG_INFO << "ZMAX " << cs->zmax;
  UPC_LOG_INFO << "MMIN " << cs->mmin;
  UPC_LOG_INFO << "MMAX " << cs->mmax;
  UPC_LOG_INFO << "YMIN " << cs->ymin;
  UPC_LOG_INFO << "YMAX " << cs->ymax;
  UPC_LOG_INFO << "BINS_Z " << cs->nz;
  UPC_LOG_INFO << "BINS_M " << cs->nm;
  UPC_LOG_INFO << "BINS_Y " << cs->ny;
  UPC_LOG_INFO << "FLUX_POINT " << cs->isPoint;
  UPC_LOG_INFO << "BREAKUP_MODE " << cs->breakupMode;
  UPC_LOG_INFO << "NON_ZERO_GAM_PT " << cs->useNonzeroGamPt;
  UPC_LOG_INFO << "USE_POLARIZED_CS " << usePolarizedCS;
  UPC_LOG_INFO << "PYTHIA_VERSION " << pythiaVersion;
  UPC_LOG_INFO << "SEED " << seed;
}

void UpcGenerator::init()
{
  if (nucProcessCS == nullptr) {
    UPC_LOG_FATAL << "UpcCrossSection was not initialized! Exiting...";
    std::_Exit(-1);
  }
  ignoreCSZ = false;
  if (procID < 11 || procID > 15 || procID == 12 || procID == 14) {
    UPC_LOG_FATAL << "PROC_ID " << procID
                  << " is not part of this engine: it covers the dilepton processes 11, 13, 15. Exiting...";
    std::_Exit(-1);
  }
  nucProcessCS->setElemProcess(procID);
  isPairProduction = true;
  auto* proc = static_cast<UpcTwoPhotonDilep*>(nucProcessCS->elemProcess);
  proc->aLep = aLep;
  const double mPart = proc->mPart;
  if (nucProcessCS->mmin < mPart * 2.) {
    UPC_LOG_WARNING << "MMIN is lower than 2 lepton masses! Setting MMIN to 2 lepton masses...";
    nucProcessCS->mmin = mPart * 2.;
  }
  nucProcessCS->numThreads = numThreads;
  if (pythiaVersion > 0) UPC_LOG_WARNING << "Decays with Pythia are not part of this build: PYTHIA_VERSION ignored";
  UPC_LOG_WARNING << "Check inputs:";
  printParameters();
  nucProcessCS->init();
  if (seed == 0) seed = (long int)std::time(nullptr);
  computeNuclXsection();
}

void UpcGenerator::computeNuclXsection()
{
  UpcCrossSection* cs = nucProcessCS;
  const int nm = cs->nm, nz = cs->nz, ny = cs->ny;
  const double mmin = cs->mmin, mmax = cs->mmax, dm = (mmax - mmin) / nm;
  const double zmin = cs->zmin, zmax = cs->zmax, dz = (zmax - zmin) / nz;
  const double ymin = cs->ymin, ymax = cs->ymax, dy = (ymax - ymin) / ny;

  std::vector<std::vector<double>> crossSectionZM, crossSectionZMPolS, crossSectionZMPolPS;
  if (usePolarizedCS) {
    cs->fillCrossSectionZM(crossSectionZMPolS, zmin, zmax, nz, mmin, mmax, nm, 1);
    cs->fillCrossSectionZM(crossSectionZMPolPS, zmin, zmax, nz, mmin, mmax, nm, 2);
  } else {
    cs->fillCrossSectionZM(crossSectionZM, zmin, zmax, nz, mmin, mmax, nm, 0);
  }
  nucCSYM.assign(ny, std::vector<double>(nm, 0.));
  if (usePolarizedCS) polCSRatio.assign(ny, std::vector<double>(nm));
  cs->calcNucCrossSectionYM(nucCSYM, polCSRatio, totCS);

  binEdgesM.resize(nm + 1);
  binEdgesZ.resize(nz + 1);
  binEdgesY.resize(ny + 1);
  for (int i = 0; i < nm + 1; ++i) binEdgesM[i] = mmin + dm * i;
  for (int i = 0; i < nz + 1; ++i) binEdgesZ[i] = zmin + dz * i;
  for (int i = 0; i < ny + 1; ++i) binEdgesY[i] = ymin + dy * i;

  // z samplers: all M rows in one GPU pass
  auto zTables = [&](const std::vector<std::vector<double>>& t) {
    std::vector<double> flat((size_t)nm * nz);
    for (int i = 0; i < nm; ++i) std::copy(t[i].begin(), t[i].end(), flat.begin() + (size_t)i * nz);
    return upc_sampler_detail::pdf_init(flat.data(), nm, nz);
  };
  auto upload = [&](const std::vector<double>& h, double** d) {
    hipOk(hipMalloc(d, sizeof(double) * h.size()), "hipMalloc");
    hipOk(hipMemcpy(*d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice), "hipMemcpy");
  };
  freeDevice();
  for (auto* s : samplersCsZ) delete s;
  for (auto* s : samplersCsSZ) delete s;
  for (auto* s : samplersCsPsZ) delete s;
  samplersCsZ.clear();
  samplersCsSZ.clear();
  samplersCsPsZ.clear();
  if (usePolarizedCS) {
    const auto ss = zTables(crossSectionZMPolS), sp = zTables(crossSectionZMPolPS);
    for (int i = 0; i < nm; ++i) {
      samplersCsSZ.push_back(new UpcSampler1D(ss.data() + (size_t)i * (nz + 1), nz, binEdgesZ, seed));
      samplersCsPsZ.push_back(new UpcSampler1D(sp.data() + (size_t)i * (nz + 1), nz, binEdgesZ, seed));
    }
    upload(ss, &d_sum_zs);
    upload(sp, &d_sum_zp);
    std::vector<double> r((size_t)ny * nm);
    for (int iy = 0; iy < ny; ++iy) std::copy(polCSRatio[iy].begin(), polCSRatio[iy].end(), r.begin() + (size_t)iy * nm);
    upload(r, &d_ratio);
  } else {
    const auto sz = zTables(crossSectionZM);
    for (int i = 0; i < nm; ++i) samplersCsZ.push_back(new UpcSampler1D(sz.data() + (size_t)i * (nz + 1), nz, binEdgesZ, seed));
    upload(sz, &d_sum_z);
  }
  delete samplerCsYM;
  samplerCsYM = new UpcSampler2D(nucCSYM, binEdgesY, binEdgesM, seed);
  upload(samplerCsYM->sum, &d_sum_ym);
  if (cs->useNonzeroGamPt) abiOk(upc_photon_pt_tables(cs->context(), cs->gtot, nullptr), "upc_photon_pt_tables");
  evP.clear();
  evNext = 0;
  evOffset = 0;
}

upc_event_opts UpcGenerator::eventOpts() const
{
  upc_event_opts o{};
  o.nonzero_gam_pt = nucProcessCS->useNonzeroGamPt ? 1 : 0;
  o.do_pt_cut = doPtCut;
  o.pt_min = minPt;
  o.do_eta_cut = doEtaCut;
  o.eta_min = minEta;
  o.eta_max = maxEta;
  return o;
}

void UpcGenerator::fetchBatch(long long n)
{
  double* dp = nullptr;
  int *dpdg = nullptr, *dpass = nullptr;
  hipOk(hipMalloc(&dp, sizeof(double) * 8 * n), "hipMalloc");
  hipOk(hipMalloc(&dpdg, sizeof(int) * 2 * n), "hipMalloc");
  hipOk(hipMalloc(&dpass, sizeof(int) * n), "hipMalloc");
  const upc_event_opts o = eventOpts();
  abiOk(upc_generate_pairs(nucProcessCS->context(), d_sum_ym, d_sum_z, d_sum_zs, d_sum_zp, d_ratio, &o,
                           (unsigned long long)seed, evOffset, n, dp, dpdg, dpass, nullptr),
        "upc_generate_pairs");
  evOffset += n;
  evP.resize(8 * n);
  evPdg.resize(2 * n);
  evPass.resize(n);
  hipOk(hipMemcpy(evP.data(), dp, sizeof(double) * 8 * n, hipMemcpyDeviceToHost), "hipMemcpy");
  hipOk(hipMemcpy(evPdg.data(), dpdg, sizeof(int) * 2 * n, hipMemcpyDeviceToHost), "hipMemcpy");
  hipOk(hipMemcpy(evPass.data(), dpass, sizeof(int) * n, hipMemcpyDeviceToHost), "hipMemcpy");
  (void)hipFree(dp);
  (void)hipFree(dpdg);
  (void)hipFree(dpass);
  evNext = 0;
}

long int UpcGenerator::generateEvent(std::vector<int>& pdgs, std::vector<int>& statuses, std::vector<int>& mothers,
                                     std::vector<UpcLorentzVector>& particles)
{
  pdgs.clear();
  statuses.clear();
  mothers.clear();
  particles.clear();
  if (evNext >= evPass.size()) fetchBatch(1 << 16);
  const size_t i = evNext++;
  if (evPass[i] != 1) return 0;  // rejected by the kinematic cuts (or not drawn)
  for (int l = 0; l < 2; ++l) {
    UpcLorentzVector v;
    v.px = evP[8 * i + 4 * l];
    v.py = evP[8 * i + 4 * l + 1];
    v.pz = evP[8 * i + 4 * l + 2];
    v.e = evP[8 * i + 4 * l + 3];
    particles.push_back(v);
    pdgs.push_back(evPdg[2 * i + l]);
    mothers.push_back(0);
    statuses.push_back(23);
  }
  return 1;
}

void UpcGenerator::writeEvent(long int evt, const std::vector<int>& pdgs, const std::vector<int>& statuses,
                              const std::vector<int>& mothers, const std::vector<UpcLorentzVector>& particles)
{
  if (!useHepMCOut || !writerHepMC) return;
  int nVertices = -1, lastMotherId = -1;
  for (int mother : mothers) {
    if (mother != lastMotherId) {
      nVertices++;
      lastMotherId = mother;
    }
  }
  writerHepMC->writeEventInfo(evt, (int)particles.size(), nVertices);
  for (size_t i = 0; i < particles.size(); ++i)
    writerHepMC->writeParticleInfo((int)i + 1, mothers[i], pdgs[i], particles[i].Px(), particles[i].Py(),
                                   particles[i].Pz(), particles[i].E(), particles[i].M(), statuses[i]);
}

void UpcGenerator::generateEvents()
{
  if (useROOTOut)
    UPC_LOG_WARNING << "ROOT output is not part of this build (no ROOT): use USE_HEPMC_OUTPUT 1 for events";
  if (useHepMCOut) {
    delete writerHepMC;
    writerHepMC = new WriterHepMC(hepmcFileName);
  }
  UPC_LOG_INFO << "Generating " << nEvents << " events...";
  std::vector<int> pdgs, statuses, mothers;
  std::vector<UpcLorentzVector> particles;
  long int rejected = 0, evt = 0, undrawn = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (evt < nEvents) {
    if (evNext >= evPass.size()) fetchBatch(std::min<long long>(1LL << 20, std::max<long long>(1024, 2 * (nEvents - evt))));
    const int st = evPass[evNext];
    if (generateEvent(pdgs, statuses, mothers, particles) == 1) {
      writeEvent(evt, pdgs, statuses, mothers, particles);
      evt++;
      if (debug <= 1 && (evt % 1000000 == 0)) UPC_LOG_INFO << "Event number: " << evt;
    } else if (st == 2) {
      rejected++;
    } else {
      undrawn++;
    }
  }
  fidCS = totCS * (double)nEvents / (double)(nEvents + rejected);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  UPC_LOG_INFO << "Event generation is finished! (" << secs << " s)";
  if (undrawn) UPC_LOG_WARNING << undrawn << " draws fell beyond the last CDF entry and were repeated";
  if ((doPtCut || doEtaCut) && nEvents > 0) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.6f", fidCS);
    UPC_LOG_INFO << "Kinematic cuts were used";
    UPC_LOG_INFO << "Number of rejected events = " << rejected;
    UPC_LOG_INFO << "Cross section with cuts = " << buf << " mb";
  }
  delete writerHepMC;
  writerHepMC = nullptr;
}
