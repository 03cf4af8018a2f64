// electronicenergy.h -- ElectronicEnergy (reference: src/common/electronicenergy.h:13-22).
#ifndef SEMIGPU_ELECTRONICENERGY_H
#define SEMIGPU_ELECTRONICENERGY_H

#include "densitymatrix.h"
#include "fockmatrix.h"
#include "hcore.h"

class ElectronicEnergy {
 public:
  ElectronicEnergy() {}
  // E_el = 1/2 sum_ij P_ij (H_ij + F_ij), eV, from the three host matrices.
  static double ComputeEnergy(const size_t& nAOs, const Hcore&, const FockMatrix&,
                              const DensityMatrix&);
};

#endif  // SEMIGPU_ELECTRONICENERGY_H
