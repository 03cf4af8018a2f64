// SCFCalculation (scfcalculation.cpp:19-203 of the reference), every stage on the GPU.
#include "scfcalculation.h"

#include <chrono>
#include <cmath>
#include <iostream>

#include "electronicenergy.h"
#include "screenutils.h"

namespace {
using clk = std::chrono::high_resolution_clock;
long ms_since(clk::time_point t) {
  return (long)std::chrono::duration_cast<std::chrono::milliseconds>(clk::now() - t).count();
}
}  // namespace

SCFCalculation::SCFCalculation(const ListAtomicOrbitals& infoAOs, double****& ints,
                               const Hcore& hcore)
    : eigenVec(nullptr), eigenVal(nullptr), Pmatrix(nullptr), Fmatrix(nullptr),
      finalEnergy(0.0), energyConverge(false), thresholdEnergy(5.0e-3), maxSCFSteps(100),
      infoAOs_(&infoAOs), nAOs_(infoAOs.size()), all2CIntegral_(ints), hcore_(&hcore),
      printInfo(false) {}

SCFCalculation::~SCFCalculation() {}

bool SCFCalculation::AllocSCFData() {
  eigenVec = new double[nAOs_ * nAOs_]();
  eigenVal = new double[nAOs_]();
  Pmatrix = new DensityMatrix(*infoAOs_, eigenVec, nAOs_);
  if (!Pmatrix->Alloc4Matrix("Pmatrix")) return false;
  Fmatrix = new FockMatrix(*infoAOs_, *hcore_, *Pmatrix, all2CIntegral_);
  return Fmatrix->Alloc4Matrix("Fmatrix");
}

bool SCFCalculation::DeallocSCFData() {
  delete[] eigenVec;
  delete[] eigenVal;
  eigenVec = eigenVal = nullptr;
  bool ok = true;
  if (Pmatrix) ok = Pmatrix->Dealloc4Matrix() && ok;
  if (Fmatrix) ok = Fmatrix->Dealloc4Matrix() && ok;
  delete Pmatrix;
  delete Fmatrix;
  Pmatrix = nullptr;
  Fmatrix = nullptr;
  return ok;
}

void SCFCalculation::InitialGuess() {
  Pmatrix->GuessDensityMatrixSwitch(true);
  Pmatrix->ComputeMatrix();
  Pmatrix->GuessDensityMatrixSwitch(false);
  Fmatrix->ComputeMatrix();
}

int SCFCalculation::GetEigenValVecOfFmatrix() {
  sg_batch* b = infoAOs_->Batch();
  Fmatrix->Update_deviceMatrix();
  semigpu::check(sg_eigh(b), "sg_eigh");
  int status = 0;
  semigpu::check(sg_get_status(b, &status), "sg_get_status");
  semigpu::check(sg_get_eigen(b, 0, eigenVal, eigenVec), "sg_get_eigen");
  return status;
}

void SCFCalculation::ComputeSCF() {
  InitialGuess();
  double electronicEnergy = ElectronicEnergy::ComputeEnergy(nAOs_, *hcore_, *Fmatrix, *Pmatrix);
  std::cout << "SCF step : " << 0 << "  Energy = " << electronicEnergy << std::endl;
  unsigned int steps = 0;
  energyConverge = false;
  while (!energyConverge && maxSCFSteps > steps) {
    const auto start = clk::now();
    auto t = clk::now();
    if (GetEigenValVecOfFmatrix() > 0) {
      std::cout << "LAPACK: The algorithm failed to converge" << std::endl;
      break;
    }
    const long lapack = ms_since(t);
    t = clk::now();
    Pmatrix->ComputeMatrix();
    const long pmat = ms_since(t);
    t = clk::now();
    Fmatrix->ComputeMatrix();
    const long fmat = ms_since(t);
    t = clk::now();
    const double old = electronicEnergy;
    electronicEnergy = ElectronicEnergy::ComputeEnergy(nAOs_, *hcore_, *Fmatrix, *Pmatrix);
    const long ener = ms_since(t);
    energyConverge = std::fabs(old - electronicEnergy) < thresholdEnergy;  // sameReal, mymath.h
    ++steps;
    std::cout << "SCF step : " << steps << "  Energy = " << electronicEnergy
              << "  LapackTime = " << lapack << "  PmatTime = " << pmat << "  FmatTime = " << fmat
              << "  EnerTime = " << ener << "  StepTime = " << ms_since(start) << std::endl;
  }
  finalEnergy = electronicEnergy;
}
