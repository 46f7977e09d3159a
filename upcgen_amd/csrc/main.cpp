// upcgen command line (main.cpp:30-116 of the reference): -h, -debug N,
// -nthreads N, -parfile FILE; plus -form fp64|mfma|surrogate, -lumidir DIR
// and -printpars (print the parsed parameters and exit, no GPU needed).
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "UpcGenerator.h"
#include "UpcLog.h"

class InputParser
{
 public:
  InputParser(int& argc, char** argv)
  {
    for (int i = 1; i < argc; ++i) tokens.emplace_back(argv[i]);
  }
  const std::string& getCmdOption(const std::string& option) const
  {
    auto itr = std::find(tokens.begin(), tokens.end(), option);
    if (itr != tokens.end() && ++itr != tokens.end()) return *itr;
    static const std::string empty;
    return empty;
  }
  bool cmdOptionExists(const std::string& option) const
  {
    return std::find(tokens.begin(), tokens.end(), option) != tokens.end();
  }

 private:
  std::vector<std::string> tokens;
};

int main(int argc, char** argv)
{
  InputParser input(argc, argv);
  int debugLevel = 0;
  int numThreads = 1;
  std::string parFileName("parameters.in");
  if (input.cmdOptionExists("-h")) {
    std::printf(
        "Available options:\n"
        "-h         -- help message\n"
        "-debug     -- set debug level: 0=no debug messages (default),\n"
        "                               1=basic debug messages,\n"
        "                               2=print out intermediate calculation results and events info\n"
        "-nthreads  -- accepted for compatibility (the luminosity runs on the GPU)\n"
        "-parfile   -- name of input parameter file\n"
        "-form      -- luminosity form: fp64 (default), mfma, surrogate\n"
        "-lumidir   -- directory of the cached luminosity grid and its lock file\n"
        "-hepmc     -- HepMC output file name (default events.hepmc)\n"
        "-locktimeout -- seconds after which a luminosity lock file is taken over\n"
        "-printpars -- print the parameters read and exit\n");
    std::fflush(stdout);
    std::_Exit(0);
  }
  const auto dbg = input.getCmdOption("-debug");
  if (!dbg.empty()) debugLevel = std::stoi(dbg);
  const auto nth = input.getCmdOption("-nthreads");
  if (!nth.empty()) numThreads = std::stoi(nth);
  const auto pf = input.getCmdOption("-parfile");
  if (!pf.empty()) parFileName = pf;

  UPC_LOG_INFO << "Configuring the generator...";
  auto* upcGenerator = new UpcGenerator();
  upcGenerator->setDebugLevel(debugLevel);
  upcGenerator->setNumThreads(numThreads);
  upcGenerator->setParFile(parFileName);
  upcGenerator->configGeneratorFromFile();
  const auto form = input.getCmdOption("-form");
  if (form == "mfma") upcGenerator->setLumiForm(UPC_FORM_MFMA);
  if (form == "surrogate") upcGenerator->setLumiForm(UPC_FORM_SURROGATE);
  const auto dir = input.getCmdOption("-lumidir");
  if (!dir.empty()) upcGenerator->setLumiFileDirectory(dir);
  const auto lt = input.getCmdOption("-locktimeout");
  if (!lt.empty()) upcGenerator->crossSection()->lockTimeoutSec = std::stoi(lt);
  const auto hm = input.getCmdOption("-hepmc");
  if (!hm.empty()) upcGenerator->hepmcFileName = hm;
  if (input.cmdOptionExists("-printpars")) {
    upcGenerator->printParameters();
    delete upcGenerator;
    return 0;
  }
  UPC_LOG_INFO << "Initializing the generator...";
  upcGenerator->init();
  UPC_LOG_INFO << "Starting generation process...";
  upcGenerator->generateEvents();
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.17g", upcGenerator->totNuclX());
  UPC_LOG_INFO << "totCS_mb " << buf;
  std::snprintf(buf, sizeof(buf), "%.17g", upcGenerator->fidNuclX());
  UPC_LOG_INFO << "fidCS_mb " << buf;
  delete upcGenerator;
  return 0;
}
