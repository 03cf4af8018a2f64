// scfcalculation.h -- SCFCalculation (reference: src/common/scfcalculation.h:12-41).
#ifndef SEMIGPU_SCFCALCULATION_H
#define SEMIGPU_SCFCALCULATION_H

#include "atomicOrbitals.h"
#include "densitymatrix.h"
#include "fockmatrix.h"
#include "hcore.h"

class SCFCalculation {
 public:
  SCFCalculation(const ListAtomicOrbitals&, double****&, const Hcore&);
  ~SCFCalculation();
  double* eigenVec;  // column-major N x N
  double* eigenVal;
  DensityMatrix* Pmatrix;
  FockMatrix* Fmatrix;
  double finalEnergy;
  bool energyConverge;
  bool AllocSCFData();
  bool DeallocSCFData();
  // Guess, then eigensolve / density / Fock / energy on the GPU until
  // |E_k - E_(k-1)| < thresholdEnergy (5e-3 eV) or maxSCFSteps (100) steps,
  // printing one line per step like the reference.
  void ComputeSCF();
  double thresholdEnergy;
  unsigned int maxSCFSteps;

 private:
  const ListAtomicOrbitals* infoAOs_;
  size_t nAOs_;
  double**** all2CIntegral_;
  const Hcore* hcore_;
  bool printInfo;
  void InitialGuess();
  int GetEigenValVecOfFmatrix();
};

#endif  // SEMIGPU_SCFCALCULATION_H
