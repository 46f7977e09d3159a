// Drives members of the UpcCrossSection facade that the upcgen executable
// does not reach (calcPhotonFlux, calcNucCrossSectionY, getPhotonPt,
// getPairMomentum) and prints their results as one JSON object for
// tests/test_gpu_facade.py.  Built by __graft_entry__.build(); needs a GPU.
#include <cstdio>
#include <vector>

#include "UpcGenerator.h"

namespace {
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

  std::printf("{\"photon_flux\": [");
  const double ys[] = {-5.0, -2.0, 0.0, 1.5, 4.0};
  for (int i = 0; i < 5; ++i) std::printf("%s[%.17g, %.17g]", i ? ", " : "", ys[i], cs.calcPhotonFlux(3.0969, ys[i]));
  std::printf("], ");

  std::vector<std::vector<double>> csY, ratio;
  double tot = 0.;
  cs.calcNucCrossSectionY(csY, ratio, tot);
  std::printf("\"cs_y\": [");
  for (size_t i = 0; i < csY.size(); ++i) std::printf("%s%.17g", i ? ", " : "", csY[i][0]);
  std::printf("], \"ratio_y\": [");
  for (size_t i = 0; i < ratio.size(); ++i) std::printf("%s%.17g", i ? ", " : "", ratio[i][0]);
  std::printf("], \"tot_y\": %.17g, ", tot);

  // photon pt at E = 2 GeV: mean and second moment of 4000 draws
  double s1 = 0., s2 = 0.;
  const int n = 4000;
  for (int i = 0; i < n; ++i) {
    const double pt = cs.getPhotonPt(2.0);
    s1 += pt;
    s2 += pt * pt;
  }
  std::printf("\"pt_mean\": %.17g, \"pt_rms\": %.17g, \"gtot\": %.17g, ", s1 / n, std::sqrt(s2 / n), cs.gtot);

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
  std::printf("\"pair_max_dm\": %.17g, \"pair_max_dy\": %.17g, \"pair_mean_pt\": %.17g, \"pair_pt_zero\": %.17g}\n",
              dm, dy, ptsum / 500, p0.Pt());
  return 0;
}
