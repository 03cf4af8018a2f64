// Hcore (hcore.cpp:11-124 of the reference): the one-electron matrix, on the GPU.
#include "hcore.h"

Hcore::Hcore(const MNDOparameter&, const ListAtomicOrbitals& infoAOs, double**** ints)
    : BaseMatrix(infoAOs.size()), infoAOs_(&infoAOs), all2CenterIntegral_(ints),
      computed_(false) {
  Bind(infoAOs.Batch(), SG_HCORE);
}

Hcore::~Hcore() {}

void Hcore::ComputeMatrix() {
  semigpu::check(sg_hcore(batch_), "sg_hcore");
  Update_hostMatrix();
  computed_ = true;
}

double Hcore::ComputeElementMatrix(const size_t& i, const size_t& j) {
  if (!computed_) ComputeMatrix();
  return Element(i, j);
}
