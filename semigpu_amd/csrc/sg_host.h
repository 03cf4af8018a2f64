// sg_host.h -- host-side definitions shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <algorithm>
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
// Where the second group's active count lives in a batch's count buffer (ints
// from n_active; the buffer holds 64).
constexpr int kSplitCountOffset = 32;

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
  // A second stream for the SCF of a batch (its small molecules beside the
  // large ones, split_views), created right after `stream` so that the two
  // run on different hardware queues; the events that fork and join them; and
  // per group and in-flight iteration, a pinned active count and the event
  // after its copy.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int* host_counts = nullptr;  // 4 ints
  hipEvent_t ev_count[4] = {};
  bool timing = false;
  int large_min = sg::kMaxBasis + 1;  // batches created from now on
  bool graphs = true;                 // capture SCF iterations as HIP graphs
  bool overlap = true;                // two-stream iteration of split batches
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

// Eigensolve, density and Fock build of the molecules of d, on stream s.
inline void scf_solve_build(sg_context* c, const BatchDev& d, hipStream_t s) {
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
}

// The molecules of a batch in two groups by the launch order (decreasing N):
// those above kRegRows, whose reduction starts with k_tridiag (bound by HBM),
// and the rest (bound by the issue of their on-chip reduction and the other
// stages).  Each is a view of the batch over its range of the order (sel),
// with the reduction's bucket ranges shifted to it and an active count of its
// own.
inline void split_views(const BatchDev& d, BatchDev& big, BatchDev& small) {
  const int nb = d.tri_head;
  big = d;
  big.n_mol = nb;
  big.sel = d.order;
  for (int k = 1; k <= kTriBuckets; ++k) big.tri_start[k] = nb;
  big.tri_start[0] = 0;
  small = d;
  small.order = d.order + nb;
  small.sel = small.order;
  small.n_mol = d.n_mol - nb;
  small.max_n = kRegRows;  // a bound: every molecule here has N <= kRegRows
  small.tri_head = 0;
  for (int k = 0; k <= kTriBuckets; ++k) small.tri_start[k] = std::max(0, d.tri_start[k] - nb);
  small.n_active = d.n_active + kSplitCountOffset;
}

// Whether the SCF of a batch runs as two groups (split_views) on two streams:
// overlap on, molecules on both sides of kRegRows, none on the
// multi-workgroup eigensolver, no damping.
inline bool can_split(const sg_context* c, const BatchDev& d) {
  return c->overlap && c->side && d.order && d.n_large == 0 && d.damping == 0.0 &&
         d.tri_head > 0 && d.tri_head < d.n_mol;
}

// One SCF iteration of the active molecules of d (scfcalculation.cpp:82-146):
// eigensolve, density, Fock, energy, convergence bookkeeping, on stream s.
inline void scf_iteration(sg_context* c, const BatchDev& d, hipStream_t s, double threshold,
                          int max_steps, double switch_ev, int fp32_cap) {
  scf_solve_build(c, d, s);
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
