// UpcGenerator: configuration, nuclear cross section and event loop of upcgen
// (include/UpcGenerator.h:55-293) for the dilepton processes, over the GPU
// luminosity engine.  Same parameters.in keys, defaults and call sequence
// (configGeneratorFromFile -> init -> generateEvents); output as the
// reference's plain HepMC writer (ROOT is not part of this build).
#pragma once

#include <cmath>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/upc_lumi.h"
#include "UpcCrossSection.h"
#include "UpcSampler.h"

class UpcGenerator
{
 public:
  UpcGenerator();
  ~UpcGenerator();
  UpcGenerator(const UpcGenerator&) = delete;
  UpcGenerator& operator=(const UpcGenerator&) = delete;

  // process-specific parameters
  int procID{15};
  double aLep{0};

  // simulation parameters
  bool doPtCut{false};
  double minPt{0};
  bool doEtaCut{false};
  double minEta{0};
  double maxEta{0};
  bool usePolarizedCS{false};
  long int seed{0};
  long int nEvents{1000};

  void setParameterValue(const std::string& parameter, const std::string& parValue);

  static int debug;

  void configGeneratorFromFile();
  void init();
  void setLumiFileDirectory(const std::string& directory) { nucProcessCS->setLumiFileDirectory(directory); }
  void setDebugLevel(int level) { debug = level; }
  void setNumThreads(int n) { numThreads = n; }
  void setParFile(const std::string& parfilename) { parFileName = parfilename; }
  void setCollisionSystem(float sqrts, int nucl_z, int nucl_a);
  void setSeed(long int seedIn) { seed = seedIn; }
  void printParameters();

  void computeNuclXsection();
  double totNuclX() { return totCS; }
  double fidNuclX() { return fidCS; }

  long int generateEvent(std::vector<int>& pdgs, std::vector<int>& statuses, std::vector<int>& mothers,
                         std::vector<UpcLorentzVector>& particles);
  void generateEvents();

  // ---- additions of this implementation ----
  void setLumiForm(int form) { nucProcessCS->lumiForm = form; }
  UpcCrossSection* crossSection() { return nucProcessCS; }
  const std::vector<std::vector<double>>& nuclearCrossSectionYM() const { return nucCSYM; }
  std::string hepmcFileName{"events.hepmc"};

 private:
  UpcCrossSection* nucProcessCS{nullptr};
  std::string parFileName{"parameters.in"};
  bool useROOTOut{true};
  bool useHepMCOut{false};
  double totCS{0};
  double fidCS{0};
  bool ignoreCSZ{false};
  int numThreads{1};
  int pythiaVersion{-1};
  bool doFSR{false};
  bool doDecays{false};
  bool isPairProduction{false};

  std::vector<std::vector<double>> nucCSYM;
  std::vector<std::vector<double>> polCSRatio;
  std::vector<double> binEdgesM;
  std::vector<double> binEdgesZ;
  std::vector<double> binEdgesY;
  std::vector<UpcSampler1D*> samplersCsZ;
  std::vector<UpcSampler1D*> samplersCsSZ;
  std::vector<UpcSampler1D*> samplersCsPsZ;
  UpcSampler2D* samplerCsYM{nullptr};

  // device copies of the sampler tables for the batched event kernel
  double* d_sum_ym{nullptr};
  double* d_sum_z{nullptr};
  double* d_sum_zs{nullptr};
  double* d_sum_zp{nullptr};
  double* d_ratio{nullptr};
  void freeDevice();

  // prefetched event batch for generateEvent()
  std::vector<double> evP;
  std::vector<int> evPdg, evPass;
  size_t evNext{0};
  unsigned long long evOffset{0};
  void fetchBatch(long long n);
  upc_event_opts eventOpts() const;

  // plain HepMC3 ASCII writer (include/UpcGenerator.h:192-242)
  class WriterHepMC
  {
   public:
    explicit WriterHepMC(const std::string& fname)
    {
      outfile.open(fname);
      outfile << "HepMC::Version 3.02.04\n"
              << "HepMC::Asciiv3-START_EVENT_LISTING\n";
    }
    ~WriterHepMC()
    {
      outfile << "HepMC::Asciiv3-END_EVENT_LISTING\n";
      outfile.close();
    }
    void writeEventInfo(long int eventID, int nParticles, int nVertices = 0)
    {
      outfile << "E " << eventID << " " << nVertices << " " << nParticles << "\n"
              << "U GEV MM\n";
    }
    void writeParticleInfo(int id, int motherID, int pdg, double px, double py, double pz, double e, double m,
                           int status)
    {
      outfile << std::setprecision(9) << "P " << id << " " << motherID << " " << pdg << " " << px << " " << py << " "
              << pz << " " << e << " " << m << " " << status << "\n";
    }
    std::ofstream outfile;
  };
  WriterHepMC* writerHepMC{nullptr};
  void writeEvent(long int evt, const std::vector<int>& pdgs, const std::vector<int>& statuses,
                  const std::vector<int>& mothers, const std::vector<UpcLorentzVector>& particles);
};
