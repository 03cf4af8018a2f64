// semigpu_device.h -- the process-wide GPU context behind the C++ classes.
//
// The classes in this directory keep the reference's names and signatures
// (LuisAlfredoNu/semiGPU, src/common/*.h) and run every stage of the SCF path
// on the GPU through the C ABI of libsemigpu.so (include/semigpu.h).  There is
// no CPU fallback: a missing GPU or library raises std::runtime_error.
#pragma once
#include <stdexcept>
#include <string>

#include "../../include/semigpu.h"

namespace semigpu {

// Context on HIP device $SEMIGPU_DEVICE (default 0), created on first use and
// kept until the process ends.
sg_context* context();

// Throws std::runtime_error("<what>: <sg_last_error()>") when rc != SG_OK.
void check(int rc, const char* what);

}  // namespace semigpu
