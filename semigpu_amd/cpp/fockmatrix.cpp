// FockMatrix (fockmatrix.cpp:10-53 of the reference), on the GPU.
#include "fockmatrix.h"

FockMatrix::FockMatrix(const ListAtomicOrbitals& infoAOs, const Hcore& hcore,
                       const DensityMatrix& P, double****& ints)
    : BaseMatrix(infoAOs.size()), infoAOs_(&infoAOs), hcore_(&hcore), Pmatrix_(&P),
      all2CIntegral_(ints) {
  Bind(infoAOs.Batch(), SG_FOCK);
}

void FockMatrix::ComputeMatrix() {
  if (hcore_->matrixHold_)
    semigpu::check(sg_set_packed(batch_, 0, SG_HCORE, hcore_->matrixHold_), "sg_set_packed H");
  if (Pmatrix_->matrixHold_)
    semigpu::check(sg_set_packed(batch_, 0, SG_DENSITY, Pmatrix_->matrixHold_), "sg_set_packed P");
  semigpu::check(sg_fock(batch_), "sg_fock");
  Update_hostMatrix();
}

double FockMatrix::ComputeElementMatrix(const size_t& i, const size_t& j) { return Element(i, j); }
