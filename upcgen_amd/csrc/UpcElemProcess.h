// Elementary gamma-gamma processes (interface of upcgen's include/UpcElemProcess.h:26-62
// and include/UpcTwoPhotonDilep.h:29-50).  The dilepton cross sections are
// evaluated on the GPU (C ABI upc_elem_cs) with the owning context.
#pragma once

#include "../../include/upc_lumi.h"

class UpcElemProcess
{
 public:
  UpcElemProcess() = default;
  virtual ~UpcElemProcess() = default;

  double mPart{};   // mass of the produced particle
  double mDght{};   // mass of decay products
  int dghtPDG{};
  int partPDG{};
  bool isCharged{};

  virtual double calcCrossSectionY(double) { return 0.; }
  virtual double calcCrossSectionM(double) { return 0.; }
  virtual double calcCrossSectionZM(double z, double m) = 0;
  virtual double calcCrossSectionMPolS(double m) = 0;
  virtual double calcCrossSectionZMPolS(double z, double m) = 0;
  virtual double calcCrossSectionMPolPS(double m) = 0;
  virtual double calcCrossSectionZMPolPS(double z, double m) = 0;
};

// gamma gamma -> l+ l- (PROC_ID 11, 13, 15), Breit-Wheeler with anomalous
// magnetic moment terms (src/UpcTwoPhotonDilep.cpp:29-134).
class UpcTwoPhotonDilep : public UpcElemProcess
{
 public:
  explicit UpcTwoPhotonDilep(int partPDG);
  ~UpcTwoPhotonDilep() override;

  double aLep{0};  // lepton anomalous magnetic moment

  double calcCrossSectionM(double m) override;        // [nb]
  double calcCrossSectionZM(double z, double m) override;  // [GeV^-2]
  double calcCrossSectionMPolS(double m) override;    // [fm^2]
  double calcCrossSectionZMPolS(double z, double m) override;
  double calcCrossSectionMPolPS(double m) override;
  double calcCrossSectionZMPolPS(double z, double m) override;

  // the device context the cross sections are evaluated with
  void attach(upc_ctx* ctx) { ctx_ = ctx; }

 private:
  upc_ctx* ctx_{nullptr};
  double* dbuf_{nullptr};
  double eval(int flag, double m, double z);
};
