// sg_host.h -- host-side definitions shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/semigpu.h"
#include "sg_kernels.h"

namespace sg {

int fail(int code, const std::string& msg);

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      return ::sg::fail(SG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

enum { ST_TWO_CENTER = 0, ST_HCORE, ST_GUESS, ST_FOCK, ST_EIGH, ST_DENSITY, ST_ENERGY, ST_CONTROL };

int check_launch(const char* what);
bool supported_element(int z);

}  // namespace sg

struct sg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool timing = false;
  double ms[SG_N_STAGES] = {0};
  int counts[SG_N_STAGES] = {0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
};

namespace sg {

// Records HIP events around one stage on the context stream when timing is on.
struct StageTimer {
  sg_context* c;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(sg_context* ctx, int st) : c(ctx), stage(st) {
    if (c->timing) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~StageTimer() {
    if (c->timing) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({stage, {a, b}});
    }
  }
};

}  // namespace sg
