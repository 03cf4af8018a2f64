// corecorerepulsion.h -- CoreCoreRepulsion (reference: src/common/corecorerepulsion.h:13-24).
#ifndef SEMIGPU_CORECOREREPULSION_H
#define SEMIGPU_CORECOREREPULSION_H

#include "MNDO_parameters.h"
#include "atomicOrbitals.h"

class CoreCoreRepulsion {
 public:
  CoreCoreRepulsion() {}
  // E_nuc in eV from the device integrals of the molecule.
  static double ComputeRepulsion(const MNDOparameter&, const ListAtomicOrbitals&, double****);
};

#endif  // SEMIGPU_CORECOREREPULSION_H
