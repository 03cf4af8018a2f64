// fockmatrix.h -- FockMatrix (reference: src/common/fockmatrix.h:14-33).
#ifndef SEMIGPU_FOCKMATRIX_H
#define SEMIGPU_FOCKMATRIX_H

#include "atomicOrbitals.h"
#include "basematrix.h"
#include "densitymatrix.h"
#include "hcore.h"
#include "twocenterintegral.h"

class FockMatrix : public BaseMatrix {
 public:
  FockMatrix(const ListAtomicOrbitals&, const Hcore&, const DensityMatrix&, double****&);
  double ComputeElementMatrix(const size_t& i, const size_t& j);
  // F = H + G(P) from the host matrices of the Hcore and DensityMatrix objects.
  void ComputeMatrix();

 private:
  const ListAtomicOrbitals* infoAOs_;
  const Hcore* hcore_;
  const DensityMatrix* Pmatrix_;
  double**** all2CIntegral_;
};

#endif  // SEMIGPU_FOCKMATRIX_H
