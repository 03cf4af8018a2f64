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

enum {
  ST_TWO_CENTER = 0, ST_HCORE, ST_GUESS, ST_FOCK, ST_TRIDIAG, ST_BISECT, ST_INVIT, ST_BACK,
  ST_DENSITY, ST_ENERGY, ST_CONTROL, ST_TRIDIAG32, ST_TRIDIAG_HEAD
};
static_assert(ST_TRIDIAG_HEAD + 1 == SG_N_STAGES, "stage list and SG_N_STAGES disagree");

int check_launch(const char* what);
bool supported_element(int z);

inline bool valid_policy(int p) {
  return p == SG_POLICY_FP64 || p == SG_POLICY_MIXED || p == SG_POLICY_MIXED_BF16 ||
         p == SG_POLICY_MIXED_EIG32;
}
// Precision of a molecule's first SCF steps under a policy (the prec flags):
// 0 FP64, 1 FP32 Fock and density, 2 FP32 Fock and split-BF16 density.
inline int reduced_prec(int policy) {
  return policy == SG_POLICY_FP64 ? 0 : policy == SG_POLICY_MIXED_BF16 ? 2 : 1;
}

}  // namespace sg

// Streams a context keeps for the size classes of its servers (class i takes
// stream i; classes beyond these create their own).
constexpr int kCtxClassStreams = 3;

struct sg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  // Created with the context, right after `stream`: HIP maps a process's
  // streams onto its hardware queues (GPU_MAX_HW_QUEUES, 4 by default) in the
  // order they are created, and streams that share a queue run in sequence.
  // Streams made this early land on queues of their own however many streams
  // (graph launches, other libraries) the process creates later, so the
  // classes of a server overlap on the device with the default queue count.
  hipStream_t class_streams[kCtxClassStreams] = {};
  bool timing = false;
  int large_min = sg::kMaxBasis + 1;  // batches created from now on
  bool graphs = true;                 // capture SCF iterations as HIP graphs
  int tri_variant[sg::kTriBuckets] = {};  // launch shape per reduction bucket (autotuner)
  double ms[SG_N_STAGES] = {0};
  int counts[SG_N_STAGES] = {0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  // Device blocks of destroyed batches, kept for the next batches (a batch
  // takes the smallest one that fits; at most kCacheBlocks are kept), and a
  // pinned host buffer through which batch descriptors are uploaded.
  std::vector<std::pair<void*, size_t>> cache;
  void* staging = nullptr;
  size_t staging_bytes = 0;
};

namespace sg {
constexpr int kCacheBlocks = 4;
// A device block of at least `bytes` for a batch: from the context's cache
// when one fits, else hipMalloc.  *got is the block's size.
hipError_t ctx_alloc(sg_context* c, size_t bytes, void** out, size_t* got);
// Returns a batch's block to the context's cache.
void ctx_release(sg_context* c, void* p, size_t bytes);
// The context's pinned staging buffer, grown to at least `bytes`.
hipError_t ctx_staging(sg_context* c, size_t bytes, void** out);
}  // namespace sg

namespace sg {

// Records HIP events around one stage on stream s (default: the context's)
// when timing is on.
struct StageTimer {
  sg_context* c;
  int stage;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(sg_context* ctx, int st, hipStream_t stream = nullptr)
      : c(ctx), stage(st), s(stream ? stream : ctx->stream) {
    if (c->timing) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, s);
    }
  }
  ~StageTimer() {
    if (c->timing) {
      (void)hipEventRecord(b, s);
      c->pending.push_back({stage, {a, b}});
    }
  }
};

// One SCF iteration of the active molecules (scfcalculation.cpp:82-146):
// eigensolve, density, Fock, energy, convergence bookkeeping.
inline void scf_iteration(sg_context* c, const BatchDev& d, hipStream_t s, double threshold,
                          int max_steps, double switch_ev, int fp32_cap);

// The eigensolver's kernels, each timed as its own stage (the reduction as
// its head, on-chip and FP32 launches).
inline void timed_eigh(sg_context* c, const BatchDev& d, hipStream_t s, bool all) {
  if (d.tri_head > 0) {
    StageTimer t(c, ST_TRIDIAG_HEAD, s);
    launch_eigh_tridiag(d, s, all, 2);
  }
  {
    StageTimer t(c, ST_TRIDIAG, s);
    launch_eigh_tridiag(d, s, all, 0);
  }
  if (tridiag_fp32_launch(d, all)) {
    StageTimer t(c, ST_TRIDIAG32, s);
    launch_eigh_tridiag(d, s, all, 1);
  }
  {
    StageTimer t(c, ST_BISECT, s);
    launch_eigh_bisect(d, s, all);
  }
  {
    StageTimer t(c, ST_INVIT, s);
    launch_eigh_invit(d, s, all);
  }
  {
    StageTimer t(c, ST_BACK, s);
    launch_eigh_back(d, s, all);
  }
}

inline void scf_iteration(sg_context* c, const BatchDev& d, hipStream_t s, double threshold,
                          int max_steps, double switch_ev, int fp32_cap) {
  timed_eigh(c, d, s, false);
  {
    StageTimer t(c, ST_DENSITY, s);
    if (d.damping > 0.0) launch_damp_keep(d, s);
    launch_density(d, s, false);
    if (d.damping > 0.0) launch_damp_mix(d, s);
  }
  {
    StageTimer t(c, ST_FOCK, s);
    launch_fock(d, s, false);
  }
  {
    StageTimer t(c, ST_ENERGY, s);
    launch_energy_fock(d, s, d.e_new, false);
  }
  {
    StageTimer t(c, ST_CONTROL, s);
    launch_reset_count(d, s);
    launch_scf_update(d, s, threshold, max_steps, switch_ev, fp32_cap);
  }
  if (d.diis_m) {  // the next eigensolve takes the extrapolated Fock matrix
    StageTimer t(c, ST_FOCK, s);
    launch_diis(d, s);
  }
}

}  // namespace sg
