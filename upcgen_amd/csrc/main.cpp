// upcgen command line.  The flags of the reference's executable (main.cpp:
// 60-116 of the upcgen tree: -h, -debug N, -nthreads N, -parfile FILE) keep
// their names and meaning; -form, -lumidir, -hepmc, -locktimeout and
// -printpars are this build's.  Options are read in one pass over argv
// against a table: every option is "-name" or "-name value", a later
// occurrence overrides an earlier one, and an unknown option or a missing
// value is reported (the reference ignores both silently) without stopping.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

#include "UpcGenerator.h"
#include "UpcLog.h"

namespace {

struct OptionSpec {
  const char* name;
  bool takes_value;
  const char* help;
};

constexpr OptionSpec kOptions[] = {
    {"-h", false, "help message"},
    {"-debug", true, "debug level: 0 = no debug messages (default), 1 = basic debug messages,\n"
                     "             2 = intermediate results and events"},
    {"-nthreads", true, "accepted for compatibility (the luminosity runs on the GPU)"},
    {"-parfile", true, "name of the input parameter file (default parameters.in)"},
    {"-form", true, "luminosity form: fp64 (default), mfma, surrogate"},
    {"-lumidir", true, "directory of the cached luminosity grid and its lock file"},
    {"-hepmc", true, "HepMC output file name (default events.hepmc)"},
    {"-locktimeout", true, "seconds after which a luminosity lock file is taken over"},
    {"-printpars", false, "print the parameters read and exit"},
};

const OptionSpec* findOption(const std::string& arg)
{
  for (const OptionSpec& o : kOptions)
    if (arg == o.name) return &o;
  return nullptr;
}

// name -> value ("" for flags); unknown options and missing values are logged
std::map<std::string, std::string> parseArgs(int argc, char** argv)
{
  std::map<std::string, std::string> seen;
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    const OptionSpec* o = findOption(arg);
    if (!o) {
      UPC_LOG_WARNING << "Ignoring unknown command-line argument " << arg;
      continue;
    }
    if (!o->takes_value) {
      seen[arg] = "";
    } else if (i + 1 < argc) {
      seen[arg] = argv[++i];
    } else {
      UPC_LOG_WARNING << "Option " << arg << " needs a value; ignored";
    }
  }
  return seen;
}

void printHelp()
{
  std::printf("Available options:\n");
  for (const OptionSpec& o : kOptions) std::printf("%-12s -- %s\n", o.name, o.help);
  std::fflush(stdout);
}

int toInt(const std::map<std::string, std::string>& opts, const char* name, int fallback)
{
  const auto it = opts.find(name);
  if (it == opts.end()) return fallback;
  char* end = nullptr;
  const long v = std::strtol(it->second.c_str(), &end, 10);
  if (end == it->second.c_str() || *end != '\0') {
    UPC_LOG_WARNING << "Option " << name << " expects an integer, got '" << it->second << "'; using " << fallback;
    return fallback;
  }
  return (int)v;
}

}  // namespace

int main(int argc, char** argv)
{
  const auto opts = parseArgs(argc, argv);
  if (opts.count("-h")) {
    printHelp();
    std::_Exit(0);
  }
  const int debugLevel = toInt(opts, "-debug", 0);
  const int numThreads = toInt(opts, "-nthreads", 1);
  const auto pf = opts.find("-parfile");
  const std::string parFileName = pf != opts.end() ? pf->second : "parameters.in";

  UPC_LOG_INFO << "Configuring the generator...";
  auto* upcGenerator = new UpcGenerator();
  upcGenerator->setDebugLevel(debugLevel);
  upcGenerator->setNumThreads(numThreads);
  upcGenerator->setParFile(parFileName);
  upcGenerator->configGeneratorFromFile();
  if (const auto f = opts.find("-form"); f != opts.end()) {
    if (f->second == "mfma") upcGenerator->setLumiForm(UPC_FORM_MFMA);
    else if (f->second == "surrogate") upcGenerator->setLumiForm(UPC_FORM_SURROGATE);
    else if (f->second != "fp64") UPC_LOG_WARNING << "Unknown -form " << f->second << "; using fp64";
  }
  if (const auto d = opts.find("-lumidir"); d != opts.end()) upcGenerator->setLumiFileDirectory(d->second);
  if (opts.count("-locktimeout")) upcGenerator->crossSection()->lockTimeoutSec = toInt(opts, "-locktimeout", 3600);
  if (const auto h = opts.find("-hepmc"); h != opts.end()) upcGenerator->hepmcFileName = h->second;
  if (opts.count("-printpars")) {
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
