// BaseMatrix (basematrix.cpp:11-55 of the reference).
#include "basematrix.h"

BaseMatrix::BaseMatrix()
    : matrixSize_(0), matrixHold_(nullptr), array1DSize_(0), batch_(nullptr), which_(-1) {}

BaseMatrix::BaseMatrix(size_t n)
    : matrixSize_(n), matrixHold_(nullptr), array1DSize_(n * (n + 1) / 2), batch_(nullptr),
      which_(-1) {}

BaseMatrix::BaseMatrix(const BaseMatrix& o)
    : matrixSize_(o.matrixSize_), matrixHold_(o.matrixHold_), array1DSize_(o.array1DSize_),
      batch_(o.batch_), which_(o.which_) {}

void BaseMatrix::Bind(sg_batch* b, int which) {
  batch_ = b;
  which_ = which;
}

bool BaseMatrix::Alloc4Matrix(string) {
  array1DSize_ = matrixSize_ * (matrixSize_ + 1) / 2;
  delete[] matrixHold_;
  matrixHold_ = new double[array1DSize_]();
  return matrixHold_ != nullptr;
}

bool BaseMatrix::Dealloc4Matrix() {
  delete[] matrixHold_;
  matrixHold_ = nullptr;
  return true;
}

void BaseMatrix::ComputeMatrix() { Update_hostMatrix(); }

double BaseMatrix::ComputeElementMatrix(const size_t& i, const size_t& j) { return Element(i, j); }

double BaseMatrix::Element(const size_t& i, const size_t& j) const {
  const size_t r = i > j ? i : j, c = i > j ? j : i;
  return matrixHold_[r * (r + 1) / 2 + c];
}

void BaseMatrix::Update_hostMatrix() {
  if (!batch_) throw std::runtime_error("BaseMatrix: not bound to a device matrix");
  if (!matrixHold_) Alloc4Matrix("");
  semigpu::check(sg_get_packed(batch_, 0, which_, matrixHold_), "sg_get_packed");
}

void BaseMatrix::Update_deviceMatrix() {
  if (!batch_) throw std::runtime_error("BaseMatrix: not bound to a device matrix");
  if (matrixHold_) semigpu::check(sg_set_packed(batch_, 0, which_, matrixHold_), "sg_set_packed");
}
