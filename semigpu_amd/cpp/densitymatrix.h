// densitymatrix.h -- DensityMatrix (reference: src/common/densitymatrix.h:9-32).
#ifndef SEMIGPU_DENSITYMATRIX_H
#define SEMIGPU_DENSITYMATRIX_H

#include "atomicOrbitals.h"
#include "basematrix.h"

class DensityMatrix : public BaseMatrix {
 public:
  // eigenVec: column-major N x N eigenvectors (ascending eigenvalues), aliased.
  DensityMatrix(const ListAtomicOrbitals& infoAOs, const double* eigenVec, const size_t nAOs);
  double ComputeElementMatrix(const size_t& i, const size_t& j);
  void GuessDensityMatrixSwitch(bool);
  // Guess P (diagonal core charge / AO count), or P = 2 C_occ C_occ^T from eigenVec.
  void ComputeMatrix();

 private:
  size_t nAOs_;
  const double* eigenVec_;
  bool guessSwitch_;
  const ListAtomicOrbitals* infoAOs_;
};

#endif  // SEMIGPU_DENSITYMATRIX_H
