// MNDO_parameters.h -- MNDOparameter (reference: src/common/MNDO_parameters.h).
//
// The MNDO parameter set lives in the device library (constant memory of the
// integral and Hcore kernels); this type only keeps the reference's
// constructor signatures.
#ifndef SEMIGPU_MNDO_PARAMETERS_H
#define SEMIGPU_MNDO_PARAMETERS_H

class MNDOparameter {
 public:
  MNDOparameter() {}
};

#endif  // SEMIGPU_MNDO_PARAMETERS_H
