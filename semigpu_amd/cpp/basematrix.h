// basematrix.h -- BaseMatrix (reference: src/common/basematrix.h:10-40).
//
// matrixHold_ is the host copy (packed lower triangle, index r(r+1)/2 + c) of a
// matrix that lives on the GPU; ComputeMatrix() computes it there and copies it
// back, Update_deviceMatrix() pushes host edits to the device.
#ifndef SEMIGPU_BASEMATRIX_H
#define SEMIGPU_BASEMATRIX_H

#include <cstddef>
#include <string>
using std::string;

#include "semigpu_device.h"

class BaseMatrix {
 public:
  BaseMatrix();
  BaseMatrix(size_t);
  BaseMatrix(const BaseMatrix&);
  virtual ~BaseMatrix() {}
  size_t matrixSize_;
  double* matrixHold_;
  size_t array1DSize_;
  bool Alloc4Matrix(string matrixName);
  bool Dealloc4Matrix();
  virtual void ComputeMatrix();
  void Update_hostMatrix();    // device -> matrixHold_
  void Update_deviceMatrix();  // matrixHold_ -> device
  sg_batch* DeviceBatch() const { return batch_; }

 protected:
  virtual double ComputeElementMatrix(const size_t& i, const size_t& j);
  // The device matrix this object mirrors (set by the derived class).
  sg_batch* batch_;
  int which_;
  void Bind(sg_batch* b, int which);
  double Element(const size_t& i, const size_t& j) const;
};

#endif  // SEMIGPU_BASEMATRIX_H
