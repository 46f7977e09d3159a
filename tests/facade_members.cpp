// Drives members of the UpcCrossSection facade that the upcgen executable
// does not reach (calcPhotonFlux, calcNucCrossSectionY, getPhotonPt,
// getPairMomentum) and prints their results as one JSON object for
// tests/test_gpu_facade.py.  Built by __graft_entry__.build(); needs a GPU.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "UpcGenerator.h"

namespace {
// The object is built up in `out` and printed once at the end: the facade's
// log lines go to the same stdout while the members run.
std::string out;
void emit(const char* fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  out += buf;
}

// an elementary process with a known sigma(y) for calcNucCrossSectionY
struct LinearY : UpcElemProcess {
  double calcCrossSectionY(double y) override { return 1.0 + 0.1 * y + 0.01 * y * y; }
  double calcCrossSectionZM(double, double) override { return 0; }
  double calcCrossSectionMPolS(double) override { return 0; }
  double calcCrossSectionZMPolS(double, double) override { return 0; }
  double calcCrossSectionMPolPS(double) override { return 0; }
  double calcCrossSectionZMPolPS(double, double) override { return 0; }
};
}  // namespace

int main()
{
  UpcCrossSection cs;
  cs.isPoint = false;
  cs.breakupMode = 2;
  cs.nm = 3;
  cs.ny = 9;
  cs.ymin = -4.0;
  cs.ymax = 4.0;
  auto* proc = new LinearY();
  proc->mPart = 3.0969;
  cs.elemProcess = proc;

  emit("{\"photon_flux\": [");
  const double ys[] = {-5.0, -2.0, 0.0, 1.5, 4.0};
  for (int i = 0; i < 5; ++i) emit("%s[%.17g, %.17g]", i ? ", " : "", ys[i], cs.calcPhotonFlux(3.0969, ys[i]));
  emit("], ");

  std::vector<std::vector<double>> csY, ratio;
  double tot = 0.;
  cs.calcNucCrossSectionY(csY, ratio, tot);
  emit("\"cs_y\": [");
  for (size_t i = 0; i < csY.size(); ++i) emit("%s%.17g", i ? ", " : "", csY[i][0]);
  emit("], \"ratio_y\": [");
  for (size_t i = 0; i < ratio.size(); ++i) emit("%s%.17g", i ? ", " : "", ratio[i][0]);
  emit("], \"tot_y\": %.17g, ", tot);

  // photon pt at E = 2 GeV: mean and second moment of 4000 draws
  double s1 = 0., s2 = 0.;
  const int n = 4000;
  for (int i = 0; i < n; ++i) {
    const double pt = cs.getPhotonPt(2.0);
    s1 += pt;
    s2 += pt * pt;
  }
  emit("\"pt_mean\": %.17g, \"pt_rms\": %.17g, \"gtot\": %.17g, ", s1 / n, std::sqrt(s2 / n), cs.gtot);

  // pair momenta: invariant mass, rapidity and transverse momentum
  double dm = 0., dy = 0., ptsum = 0.;
  for (int i = 0; i < 500; ++i) {
    UpcLorentzVector p;
    cs.getPairMomentum(10.0, 1.2, p);
    dm = std::fmax(dm, std::fabs(p.M() - 10.0));
    const double y = 0.5 * std::log((p.E() + p.Pz()) / (p.E() - p.Pz()));
    dy = std::fmax(dy, std::fabs(y - 1.2));
    ptsum += p.Pt();
  }
  cs.useNonzeroGamPt = false;
  UpcLorentzVector p0;
  cs.getPairMomentum(10.0, 1.2, p0);
  emit("\"pair_max_dm\": %.17g, \"pair_max_dy\": %.17g, \"pair_mean_pt\": %.17g, \"pair_pt_zero\": %.17g}\n",
              dm, dy, ptsum / 500, p0.Pt());
  std::fflush(stdout);
  std::fputs(out.c_str(), stdout);
  return 0;
}
