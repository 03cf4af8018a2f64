// hcore.h -- Hcore (reference: src/common/hcore.h:16-45).
#ifndef SEMIGPU_HCORE_H
#define SEMIGPU_HCORE_H

#include "MNDO_parameters.h"
#include "atomicOrbitals.h"
#include "basematrix.h"

class Hcore : public BaseMatrix {
 public:
  Hcore(const MNDOparameter&, const ListAtomicOrbitals&, double****);
  ~Hcore();
  // Element (i, j) of the matrix ComputeMatrix produced (computed on first use).
  double ComputeElementMatrix(const size_t& i, const size_t& j);
  void ComputeMatrix();

 private:
  const ListAtomicOrbitals* infoAOs_;
  double**** all2CenterIntegral_;
  bool computed_;
};

#endif  // SEMIGPU_HCORE_H
