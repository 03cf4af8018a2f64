// DensityMatrix (densitymatrix.cpp:9-47 of the reference), on the GPU.
#include "densitymatrix.h"

DensityMatrix::DensityMatrix(const ListAtomicOrbitals& infoAOs, const double* eigenVec,
                             const size_t nAOs)
    : BaseMatrix(nAOs), nAOs_(nAOs), eigenVec_(eigenVec), guessSwitch_(false),
      infoAOs_(&infoAOs) {
  Bind(infoAOs.Batch(), SG_DENSITY);
}

void DensityMatrix::GuessDensityMatrixSwitch(bool on) { guessSwitch_ = on; }

void DensityMatrix::ComputeMatrix() {
  if (guessSwitch_) {
    semigpu::check(sg_guess_density(batch_), "sg_guess_density");
  } else {
    semigpu::check(sg_set_eigen(batch_, 0, nullptr, eigenVec_), "sg_set_eigen");
    semigpu::check(sg_density(batch_), "sg_density");
  }
  Update_hostMatrix();
}

double DensityMatrix::ComputeElementMatrix(const size_t& i, const size_t& j) { return Element(i, j); }
