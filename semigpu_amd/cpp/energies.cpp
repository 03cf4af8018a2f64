// ElectronicEnergy::ComputeEnergy (electronicenergy.cpp:12-36) and
// CoreCoreRepulsion::ComputeRepulsion (corecorerepulsion.cpp:12-41), on the GPU.
#include "corecorerepulsion.h"
#include "electronicenergy.h"

double ElectronicEnergy::ComputeEnergy(const size_t&, const Hcore& H, const FockMatrix& F,
                                       const DensityMatrix& P) {
  // the three matrices belong to one molecule: push the host copies, sum on the device
  sg_batch* b = H.DeviceBatch();
  if (!b || b != F.DeviceBatch() || b != P.DeviceBatch())
    throw std::runtime_error("ComputeEnergy: matrices of different molecules");
  if (H.matrixHold_) semigpu::check(sg_set_packed(b, 0, SG_HCORE, H.matrixHold_), "sg_set_packed H");
  if (F.matrixHold_) semigpu::check(sg_set_packed(b, 0, SG_FOCK, F.matrixHold_), "sg_set_packed F");
  if (P.matrixHold_) semigpu::check(sg_set_packed(b, 0, SG_DENSITY, P.matrixHold_), "sg_set_packed P");
  double e = 0.0;
  semigpu::check(sg_energy(b, &e), "sg_energy");
  return e;
}

double CoreCoreRepulsion::ComputeRepulsion(const MNDOparameter&, const ListAtomicOrbitals& infoAOs,
                                           double****) {
  double e = 0.0;
  semigpu::check(sg_core_core(infoAOs.Batch(), &e), "sg_core_core");
  return e;
}
