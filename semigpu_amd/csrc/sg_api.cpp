// sg_api.cpp -- host implementation of the C ABI in include/semigpu.h.
//
// A sg_batch owns one device allocation that holds every molecule's atoms,
// pair blocks, matrices and SCF state (layout in sg_kernels.h).  The SCF loop
// runs on the device; the host only reads back the count of molecules that are
// still iterating after each step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "sg_host.h"

using namespace sg;

static thread_local std::string g_err;

namespace sg {

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return SG_OK;
}

bool supported_element(int z) { return z == 1 || z == 6 || z == 7 || z == 8; }

}  // namespace sg

struct sg_batch {
  sg_context* ctx = nullptr;
  BatchDev d;
  void* dev_mem = nullptr;
  std::vector<MolDesc> mols;
  std::vector<AtomDesc> atoms;
  std::vector<int64_t> blk_count;  // doubles of pair blocks per molecule
  bool have_ints = false, have_hcore = false, have_ecc = false, have_fp32 = false;
  bool have_scf = false;  // an SCF has run: P is the converged density
  std::vector<double> e_cc;
  std::vector<int> large_idx;       // molecules on the multi-workgroup eigensolver
  std::vector<MolDesc> large_mols;  // their descriptors
  size_t dev_bytes = 0;             // size of dev_mem (returned to the context's cache)
  // captured SCF iteration (of the first group when the SCF runs as two
  // groups, split_views) and of the second group
  hipGraphExec_t iter_exec = nullptr, iter_exec2 = nullptr;
  void* hist_mem = nullptr;             // per-iteration energies (sg_scf_record_energies)
  double g_threshold = -1.0, g_switch = -1.0;
  int g_max_steps = -1, g_mixed = -1;
  double g_damping = -1.0;
  void* diis_mem = nullptr;  // DIIS histories (sg_scf_set_diis)
};

static void drop_graphs(sg_batch* b) {
  if (b->iter_exec) (void)hipGraphExecDestroy(b->iter_exec);
  if (b->iter_exec2) (void)hipGraphExecDestroy(b->iter_exec2);
  b->iter_exec = b->iter_exec2 = nullptr;
}

// Where atom i's row starts in each list of a batch: the all-pairs list, the
// type-sorted lists (-1 where the atom's kind has no row there), the
// one-center tasks, and the atom's slot in its molecule's partner order
// (heavy atoms, then hydrogens).
struct ListOffsets {
  int64_t all, xx, hx, hh, diag;
  int64_t partner;
};

// One wavefront per atom (row a of its molecule): the pairs (a, j < a) into
// the all-pairs list and the list of their kind, in column order (a heavy
// column j is the hv_before(j)-th heavy atom before a, a hydrogen column the
// (j - hv_before(j))-th hydrogen), the atom's one-center tasks, and its copy
// in the partner order of the Fock build.
__global__ void __launch_bounds__(256) k_build_lists(int n_atoms, const MolDesc* __restrict__ mols,
                                                     const AtomDesc* __restrict__ atoms,
                                                     const ListOffsets* __restrict__ offs,
                                                     int2* __restrict__ pall, int2* __restrict__ pxx,
                                                     int2* __restrict__ phx, int2* __restrict__ phh,
                                                     int2* __restrict__ diag,
                                                     AtomDesc* __restrict__ partners) {
  const int gi = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (gi >= n_atoms) return;
  const AtomDesc A = atoms[gi];
  const ListOffsets o = offs[gi];
  const int atom0 = mols[A.mol].atom0;
  const bool heavy = A.zn > 1;
  for (int j = lane; j < A.local; j += 64) {
    const int gj = atom0 + j;
    const AtomDesc B = atoms[gj];
    const int2 pr = make_int2(gi, gj);
    pall[o.all + j] = pr;
    const bool hj = B.zn > 1;
    const int hb = j - B.hv_before;
    if (heavy && hj) pxx[o.xx + B.hv_before] = pr;
    else if (heavy) phx[o.hx + hb] = pr;
    else if (hj) phx[o.hx + B.hv_before] = pr;
    else phh[o.hh + hb] = pr;
  }
  if (lane < (heavy ? 10 : 1)) diag[o.diag + lane] = make_int2(gi, lane);
  if (lane == 0) partners[o.partner] = A;
}

// Workspace of sg_eigh_noalloc: one file-scope device array per GPU, sized for
// the largest matrix that entry point takes (kNoAllocMaxN): the molecule
// descriptor, order and status words, the Gershgorin bound, d, e, tau and the
// eigensolver's 5 N^2 work matrix.  Host code never uses the array's name as
// an address: sg_ctx_create looks up its device address per GPU
// (hipGetSymbolAddress) and that address is what the kernels get.
constexpr int kNoAllocMaxN = 512;
constexpr size_t kNoAllocHead = 1024;  // bytes: descriptor, order, status, bound
constexpr size_t kNoAllocDoubles =
    kNoAllocHead / 8 + 3 * (size_t)kNoAllocMaxN + 5 * (size_t)kNoAllocMaxN * kNoAllocMaxN;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
static void* g_scratch_ptr[64];  // device address of `scratch` per device ordinal

// Writes the single-molecule descriptor, order and status of an
// sg_eigh_noalloc call into its workspace (a kernel rather than copies, so
// that nothing in the call allocates staging memory or records a memset).
__global__ void k_noalloc_setup(MolDesc* __restrict__ mol, int* __restrict__ order,
                                int* __restrict__ status, int n) {
  if (threadIdx.x == 0) {
    MolDesc M;
    M.atom0 = 0;
    M.n_atoms = 0;
    M.n_ao = n;
    M.n_occ = n / 2;
    M.mat0 = 0;
    M.blk0 = 0;
    M.vec0 = 0;
    *mol = M;
    *order = 0;
    *status = 0;
  }
}

// Plain copy, 16 bytes per lane per step: reads and writes exactly n16 * 16 bytes.
__global__ void k_calibration_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                   int64_t n16) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

namespace sg {

hipError_t ctx_alloc(sg_context* c, size_t bytes, void** out, size_t* got) {
  int best = -1;
  for (int i = 0; i < (int)c->cache.size(); ++i)
    if (c->cache[i].second >= bytes && (best < 0 || c->cache[i].second < c->cache[best].second))
      best = i;
  if (best >= 0) {
    *out = c->cache[best].first;
    *got = c->cache[best].second;
    c->cache.erase(c->cache.begin() + best);
    return hipSuccess;
  }
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess && !c->cache.empty()) {  // give the cached blocks back and try once more
    (void)hipGetLastError();
    for (auto& blk : c->cache) (void)hipFree(blk.first);
    c->cache.clear();
    e = hipMalloc(out, bytes);
  }
  *got = e == hipSuccess ? bytes : 0;
  return e;
}

void ctx_release(sg_context* c, void* p, size_t bytes) {
  if (!p) return;
  c->cache.push_back({p, bytes});
  if ((int)c->cache.size() > kCacheBlocks) {  // drop the smallest
    int small = 0;
    for (int i = 1; i < (int)c->cache.size(); ++i)
      if (c->cache[i].second < c->cache[small].second) small = i;
    (void)hipFree(c->cache[small].first);
    c->cache.erase(c->cache.begin() + small);
  }
}

hipError_t ctx_staging(sg_context* c, size_t bytes, void** out) {
  if (bytes > c->staging_bytes) {
    const size_t want = std::max(bytes, c->staging_bytes * 2);
    if (c->staging) (void)hipHostFree(c->staging);
    c->staging = nullptr;
    c->staging_bytes = 0;
    hipError_t e = hipHostMalloc(&c->staging, want, hipHostMallocDefault);
    if (e != hipSuccess) {
      c->staging = nullptr;
      return e;
    }
    c->staging_bytes = want;
  }
  *out = c->staging;
  return hipSuccess;
}

}  // namespace sg

// Synchronizes and destroys a context's streams, events and pinned counts
// (those created so far).
static void ctx_free_handles(sg_context* c) {
  for (hipStream_t cs : c->class_streams)
    if (cs) {
      (void)hipStreamSynchronize(cs);
      (void)hipStreamDestroy(cs);
    }
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipStreamDestroy(c->side);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  for (hipEvent_t ev : c->ev_count)
    if (ev) (void)hipEventDestroy(ev);
  if (c->host_counts) (void)hipHostFree(c->host_counts);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
}

// Captures one SCF iteration of the molecules of d on stream s as a graph.
static int capture_iteration(sg_context* c, const BatchDev& d, hipStream_t s, double threshold,
                             int max_steps, double switch_ev, int fp32_cap, hipGraphExec_t* out) {
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  scf_iteration(c, d, s, threshold, max_steps, switch_ev, fp32_cap);
  HIP_TRY(hipStreamEndCapture(s, &g));
  int r;
  if ((r = check_launch("scf step capture"))) {
    (void)hipGraphDestroy(g);
    return r;
  }
  hipError_t ge = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ge != hipSuccess) {
    *out = nullptr;
    return fail(SG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ge));
  }
  return SG_OK;
}

extern "C" {

const char* sg_last_error(void) { return g_err.c_str(); }
int sg_version(void) { return 100; }

int sg_ctx_create(int device, sg_context** out) {
  if (!out) return fail(SG_ERR_INVALID, "null output handle");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
    return fail(SG_ERR_NO_DEVICE, "no HIP device " + std::to_string(device));
  HIP_TRY(hipSetDevice(device));
  if (device < 64 && !g_scratch_ptr[device]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
  }
  auto* c = new sg_context();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  for (int i = 0; i < kCtxClassStreams && e == hipSuccess; ++i)
    e = hipStreamCreateWithFlags(&c->class_streams[i], hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  for (hipEvent_t& ev : c->ev_count)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_counts, 4 * sizeof(int));
  if (e != hipSuccess) {
    ctx_free_handles(c);
    delete c;
    return fail(SG_ERR_HIP, std::string("context set-up: ") + hipGetErrorString(e));
  }
  *out = c;
  return SG_OK;
}

int sg_ctx_destroy(sg_context* c) {
  if (!c) return SG_OK;
  (void)hipSetDevice(c->device);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& blk : c->cache) (void)hipFree(blk.first);
  if (c->staging) (void)hipHostFree(c->staging);
  ctx_free_handles(c);
  delete c;
  return SG_OK;
}

int sg_ctx_set_large_basis(sg_context* c, int n_min) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  if (n_min < kMinLarge)
    return fail(SG_ERR_INVALID, "large-basis threshold below " + std::to_string(kMinLarge));
  c->large_min = n_min;
  return SG_OK;
}

int sg_ctx_autotune(sg_context* c, int molecules, double* ms, int* choice) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  if (molecules < 1 || molecules > 4096)
    return fail(SG_ERR_INVALID, "sg_ctx_autotune: molecules must lie in [1, 4096]");
  HIP_TRY(hipSetDevice(c->device));
  double t[kTriBuckets * kTriCandidates];
  int ch[kTriBuckets];
  HIP_TRY(tri_autotune(c->stream, molecules, 2, t, ch));
  for (int k = 0; k < kTriBuckets; ++k) c->tri_variant[k] = ch[k];
  if (ms) std::copy(t, t + kTriBuckets * kTriCandidates, ms);
  if (choice) std::copy(ch, ch + kTriBuckets, choice);
  return SG_OK;
}

int sg_ctx_set_tri_variant(sg_context* c, int bucket, int variant) {
  if (!c || bucket < 0 || bucket >= kTriBuckets) return fail(SG_ERR_INVALID, "bad bucket");
  int count, nw[kTriCandidates], nr[kTriCandidates];
  tri_candidates(bucket, &count, nw, nr);
  if (variant < 0 || variant >= count) return fail(SG_ERR_INVALID, "bad launch shape for the bucket");
  c->tri_variant[bucket] = variant;
  return SG_OK;
}

int sg_ctx_tri_shapes(const sg_context* c, int* variant, int* count, int* nw, int* nr) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  for (int k = 0; k < kTriBuckets; ++k) {
    if (variant) variant[k] = c->tri_variant[k];
    int cnt, w[kTriCandidates], r[kTriCandidates];
    tri_candidates(k, &cnt, w, r);
    if (count) count[k] = cnt;
    for (int j = 0; j < kTriCandidates; ++j) {
      if (nw) nw[k * kTriCandidates + j] = w[j];
      if (nr) nr[k * kTriCandidates + j] = r[j];
    }
  }
  return SG_OK;
}

int sg_ctx_set_graphs(sg_context* c, int on) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  c->graphs = on != 0;
  return SG_OK;
}

int sg_ctx_set_overlap(sg_context* c, int on) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  c->overlap = on != 0;
  return SG_OK;
}

int sg_calibration_copy(sg_context* c, long long bytes) {
  if (!c || bytes <= 0 || bytes % 16) return fail(SG_ERR_INVALID, "sg_calibration_copy: bytes must be a positive multiple of 16");
  HIP_TRY(hipSetDevice(c->device));
  void *src = nullptr, *dst = nullptr;
  HIP_TRY(hipMalloc(&src, (size_t)bytes));
  hipError_t e = hipMalloc(&dst, (size_t)bytes);
  if (e == hipSuccess) e = hipMemsetAsync(src, 0, (size_t)bytes, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) {
    k_calibration_copy<<<4096, 256, 0, c->stream>>>((const double2*)src, (double2*)dst, bytes / 16);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (e != hipSuccess) return fail(SG_ERR_HIP, std::string("sg_calibration_copy: ") + hipGetErrorString(e));
  return SG_OK;
}

int sg_timing_enable(sg_context* c, int on) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  c->timing = on != 0;
  return SG_OK;
}

int sg_timing_read(sg_context* c, double* ms, int* counts) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float t = 0.f;
    (void)hipEventElapsedTime(&t, p.second.first, p.second.second);
    c->ms[p.first] += t;
    c->counts[p.first] += 1;
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  c->pending.clear();
  for (int i = 0; i < SG_N_STAGES; ++i) {
    if (ms) ms[i] = c->ms[i];
    if (counts) counts[i] = c->counts[i];
  }
  return SG_OK;
}

int sg_timing_reset(sg_context* c) {
  if (!c) return fail(SG_ERR_INVALID, "null context");
  int r = sg_timing_read(c, nullptr, nullptr);
  for (int i = 0; i < SG_N_STAGES; ++i) {
    c->ms[i] = 0;
    c->counts[i] = 0;
  }
  return r;
}

static int batch_create(sg_context* ctx, int n_mol, const int* atom_offsets, const int* Z,
                        const double* xyz_ang, double to_bohr, sg_batch** out);

int sg_batch_create(sg_context* ctx, int n_mol, const int* atom_offsets, const int* Z,
                    const double* xyz_ang, sg_batch** out) {
  // AtomicOrbital::SetCoordinates: Angstrom -> bohr (atomicOrbitals.cpp:40-44)
  return batch_create(ctx, n_mol, atom_offsets, Z, xyz_ang, 1.8897259886, out);
}

int sg_batch_create_bohr(sg_context* ctx, int n_mol, const int* atom_offsets, const int* Z,
                         const double* xyz_bohr, sg_batch** out) {
  return batch_create(ctx, n_mol, atom_offsets, Z, xyz_bohr, 1.0, out);
}

static int batch_create(sg_context* ctx, int n_mol, const int* atom_offsets, const int* Z,
                        const double* xyz_ang, double to_bohr, sg_batch** out) {
  if (!ctx || !out || n_mol <= 0 || !atom_offsets || !Z || !xyz_ang)
    return fail(SG_ERR_INVALID, "sg_batch_create: bad arguments");
  if (atom_offsets[0] != 0) return fail(SG_ERR_INVALID, "atom_offsets[0] must be 0");
  for (int m = 0; m < n_mol; ++m)
    if (atom_offsets[m + 1] <= atom_offsets[m])
      return fail(SG_ERR_INVALID, "molecule " + std::to_string(m) + " has no atoms");
  const int n_atoms = atom_offsets[n_mol];
  for (int i = 0; i < n_atoms; ++i) {
    if (!supported_element(Z[i]))
      return fail(SG_ERR_UNSUPPORTED, "atom " + std::to_string(i) + ": element Z=" +
                                          std::to_string(Z[i]) + " has no MNDO parameters here");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(xyz_ang[3 * i + k]))
        return fail(SG_ERR_INVALID, "atom " + std::to_string(i) + " has a non-finite coordinate");
  }
  auto* b = new sg_batch();
  b->ctx = ctx;
  b->mols.resize(n_mol);
  b->atoms.resize(n_atoms);
  b->blk_count.resize(n_mol);
  // Host work is O(atoms): descriptors and, per atom, where its row of each
  // pair list starts; the lists themselves are written on the device
  // (k_build_lists), in the order of a host loop over rows a and columns
  // j < a of each molecule.
  std::vector<ListOffsets> offs(n_atoms);
  int64_t mat0 = 0, vec0 = 0, blk0 = 0;
  int64_t n_all = 0, n_xx = 0, n_hx = 0, n_hh = 0, n_diag = 0;
  int max_n = 0;
  for (int m = 0; m < n_mol; ++m) {
    MolDesc& M = b->mols[m];
    M.atom0 = atom_offsets[m];
    M.n_atoms = atom_offsets[m + 1] - atom_offsets[m];
    int ao = 0, ne = 0, hv = 0, n_heavy = 0;
    for (int a = 0; a < M.n_atoms; ++a) n_heavy += Z[M.atom0 + a] > 1;
    int64_t row = 0, hx_run = 0;
    for (int a = 0; a < M.n_atoms; ++a) {
      const int gi = M.atom0 + a;
      const bool heavy = Z[gi] > 1;
      const int hb = a - hv;  // hydrogens before a
      AtomDesc& A = b->atoms[gi];
      A.zn = Z[gi];
      A.x = xyz_ang[3 * gi + 0] * to_bohr;
      A.y = xyz_ang[3 * gi + 1] * to_bohr;
      A.z = xyz_ang[3 * gi + 2] * to_bohr;
      A.mol = m;
      A.local = a;
      A.ao0 = ao;
      A.hv_before = hv;
      A.pad = 0;
      A.row0 = row;
      // blocks (a, j <= a): 100 per heavy pair, 10 per mixed, 1 per H-H
      row += heavy ? 100 * (int64_t)(hv + 1) + 10 * (int64_t)hb : 10 * (int64_t)hv + hb + 1;
      ListOffsets& o = offs[gi];
      o.all = n_all + (int64_t)a * (a - 1) / 2;
      o.xx = heavy ? n_xx + (int64_t)hv * (hv - 1) / 2 : -1;
      o.hh = heavy ? -1 : n_hh + (int64_t)hb * (hb - 1) / 2;
      o.hx = n_hx + hx_run;
      hx_run += heavy ? hb : hv;
      o.diag = n_diag;
      o.partner = M.atom0 + (heavy ? hv : n_heavy + hb);
      n_diag += heavy ? 10 : 1;
      ao += n_ao_of(A.zn);
      ne += core_charge(A.zn);
      if (heavy) ++hv;
    }
    const int64_t nh = M.n_atoms - n_heavy;
    n_all += (int64_t)M.n_atoms * (M.n_atoms - 1) / 2;
    n_xx += (int64_t)n_heavy * (n_heavy - 1) / 2;
    n_hx += (int64_t)n_heavy * nh;
    n_hh += nh * (nh - 1) / 2;
    M.n_ao = ao;
    M.n_occ = ne / 2;
    M.mat0 = mat0;
    M.vec0 = vec0;
    M.blk0 = blk0;
    b->blk_count[m] = row;
    mat0 += (int64_t)ao * ao;
    vec0 += ao;
    blk0 += row;
    max_n = std::max(max_n, ao);
  }
  if (max_n > kMaxLarge) {
    delete b;
    return fail(SG_ERR_INVALID, "molecules above " + std::to_string(kMaxLarge) +
                                    " basis functions are not supported");
  }
  if (n_all > INT32_MAX) {
    delete b;
    return fail(SG_ERR_INVALID, "more than 2^31 - 1 atom pairs in one batch");
  }
  BatchDev& d = b->d;
  d.n_mol = n_mol;
  d.n_atoms = n_atoms;
  d.n_mat = mat0;
  d.n_vec = vec0;
  d.n_blocks = blk0;
  d.max_n = max_n;
  set_launch_shapes(d, b->mols.data(), n_mol);
  d.n_xx = (int)n_xx;
  d.n_hx = (int)n_hx;
  d.n_hh = (int)n_hh;
  d.n_pairs_all = (int)n_all;
  d.n_diag_tasks = (int)n_diag;
  d.large_min = std::min(ctx->large_min, kMaxBasis + 1);
  for (int k = 0; k < kTriBuckets; ++k) d.tri_variant[k] = ctx->tri_variant[k];

  // one allocation, 256-byte aligned sub-buffers
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_mols = take(sizeof(MolDesc) * n_mol), o_atoms = take(sizeof(AtomDesc) * n_atoms),
               o_partners = take(sizeof(AtomDesc) * n_atoms);
  const size_t o_xx = take(sizeof(int2) * n_xx), o_hx = take(sizeof(int2) * n_hx),
               o_hh = take(sizeof(int2) * n_hh), o_all = take(sizeof(int2) * n_all),
               o_diag = take(sizeof(int2) * n_diag), o_offs = take(sizeof(ListOffsets) * n_atoms);
  const size_t o_blk = take(sizeof(double) * blk0), o_blk32 = take(sizeof(float) * blk0);
  const size_t o_H = take(sizeof(double) * mat0), o_F = take(sizeof(double) * mat0),
               o_P = take(sizeof(double) * mat0), o_C = take(sizeof(double) * mat0),
               o_W = take(sizeof(double) * 5 * mat0);
  const size_t o_eig = take(sizeof(double) * vec0), o_dlam = take(sizeof(double) * vec0),
               o_tau = take(sizeof(double) * vec0),
               o_dT = take(sizeof(double) * vec0), o_eT = take(sizeof(double) * vec0),
               o_tn = take(sizeof(double) * n_mol);
  const size_t o_ecc = take(sizeof(double) * n_mol), o_eel = take(sizeof(double) * n_mol),
               o_enew = take(sizeof(double) * n_mol), o_eatom = take(sizeof(double) * n_atoms);
  const size_t o_steps = take(sizeof(int) * n_mol), o_act = take(sizeof(int) * n_mol),
               o_conv = take(sizeof(int) * n_mol), o_stat = take(sizeof(int) * n_mol),
               o_nact = take(sizeof(int) * 64), o_prec = take(sizeof(int) * n_mol),
               o_pprec = take(sizeof(int) * n_mol), o_order = take(sizeof(int) * n_mol),
               o_s32 = take(sizeof(int) * n_mol), o_dep = take(sizeof(double) * n_mol),
               o_wdiag = take(sizeof(double) * 10 * (size_t)n_atoms);
  (void)hipSetDevice(ctx->device);
  hipError_t e = ctx_alloc(ctx, off, &b->dev_mem, &b->dev_bytes);
  if (e != hipSuccess) {
    delete b;
    return fail(SG_ERR_HIP, std::string("hipMalloc(") + std::to_string(off) + "): " +
                                hipGetErrorString(e));
  }
  char* base = (char*)b->dev_mem;
  d.mols = (MolDesc*)(base + o_mols);
  d.atoms = (AtomDesc*)(base + o_atoms);
  d.partners = (AtomDesc*)(base + o_partners);
  d.pairs_xx = (int2*)(base + o_xx);
  d.pairs_hx = (int2*)(base + o_hx);
  d.pairs_hh = (int2*)(base + o_hh);
  d.pairs_all = (int2*)(base + o_all);
  d.diag_tasks = (int2*)(base + o_diag);
  d.blocks = (double*)(base + o_blk);
  d.blocks32 = (float*)(base + o_blk32);
  d.H = (double*)(base + o_H);
  d.F = (double*)(base + o_F);
  d.P = (double*)(base + o_P);
  d.C = (double*)(base + o_C);
  d.W = (double*)(base + o_W);
  d.eig = (double*)(base + o_eig);
  d.dlam = (double*)(base + o_dlam);
  d.tau = (double*)(base + o_tau);
  d.dT = (double*)(base + o_dT);
  d.eT = (double*)(base + o_eT);
  d.tnorm = (double*)(base + o_tn);
  d.e_cc = (double*)(base + o_ecc);
  d.e_el = (double*)(base + o_eel);
  d.e_new = (double*)(base + o_enew);
  d.e_atom = (double*)(base + o_eatom);
  d.steps = (int*)(base + o_steps);
  d.active = (int*)(base + o_act);
  d.converged = (int*)(base + o_conv);
  d.status = (int*)(base + o_stat);
  d.n_active = (int*)(base + o_nact);
  d.prec = (int*)(base + o_prec);
  d.prev_prec = (int*)(base + o_pprec);
  d.steps32 = (int*)(base + o_s32);
  d.de_prev = (double*)(base + o_dep);
  d.wdiag = (double*)(base + o_wdiag);
  int* order = (int*)(base + o_order);
  d.order = order;
  hipStream_t s = ctx->stream;
  // largest molecules first in the per-molecule launches
  std::vector<int> ord(n_mol);
  for (int m = 0; m < n_mol; ++m) ord[m] = m;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int x, int y) { return b->mols[x].n_ao > b->mols[y].n_ao; });
  // descriptors, list offsets and order through the pinned staging buffer:
  // one host copy, then one DMA per array
  const size_t sz_mols = sizeof(MolDesc) * n_mol, sz_atoms = sizeof(AtomDesc) * n_atoms,
               sz_offs = sizeof(ListOffsets) * n_atoms, sz_ord = sizeof(int) * n_mol;
  void* stage = nullptr;
  e = ctx_staging(ctx, sz_mols + sz_atoms + sz_offs + sz_ord, &stage);
  if (e == hipSuccess) {
    char* h = (char*)stage;
    memcpy(h, b->mols.data(), sz_mols);
    memcpy(h + sz_mols, b->atoms.data(), sz_atoms);
    memcpy(h + sz_mols + sz_atoms, offs.data(), sz_offs);
    memcpy(h + sz_mols + sz_atoms + sz_offs, ord.data(), sz_ord);
    e = hipMemcpyAsync(d.mols, h, sz_mols, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d.atoms, h + sz_mols, sz_atoms, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(base + o_offs, h + sz_mols + sz_atoms, sz_offs, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(order, h + sz_mols + sz_atoms + sz_offs, sz_ord, hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) {
    k_build_lists<<<(n_atoms + 3) / 4, 256, 0, s>>>(n_atoms, d.mols, d.atoms,
                                                    (const ListOffsets*)(base + o_offs), d.pairs_all,
                                                    d.pairs_xx, d.pairs_hx, d.pairs_hh,
                                                    d.diag_tasks, d.partners);
    e = hipGetLastError();
  }
  // the molecules at or above large_min lead the order
  for (int i = 0; i < n_mol; ++i) {
    const MolDesc& M = b->mols[ord[i]];
    if (M.n_ao >= d.large_min) {
      b->large_idx.push_back(ord[i]);
      b->large_mols.push_back(M);
    } else {
      d.max_n_small = std::max(d.max_n_small, M.n_ao);
    }
  }
  d.n_large = (int)b->large_idx.size();
  d.large_idx = b->large_idx.data();
  d.large_mols = b->large_mols.data();
  {
    std::vector<int> nsmall;
    nsmall.reserve(n_mol - d.n_large);
    for (int i = d.n_large; i < n_mol; ++i) nsmall.push_back(b->mols[ord[i]].n_ao);
    set_tri_buckets(d, nsmall.data(), (int)nsmall.size());
  }
  for (double* z : {d.H, d.F, d.P, d.C})
    if (e == hipSuccess) e = hipMemsetAsync(z, 0, sizeof(double) * mat0, s);
  if (e == hipSuccess) e = hipMemsetAsync(d.status, 0, sizeof(int) * n_mol, s);
  if (e == hipSuccess) e = hipMemsetAsync(d.active, 0, sizeof(int) * n_mol, s);
  if (e == hipSuccess) e = hipMemsetAsync(d.prec, 0, sizeof(int) * n_mol, s);
  if (e == hipSuccess) e = hipMemsetAsync(d.steps32, 0, sizeof(int) * n_mol, s);
  if (e == hipSuccess) e = hipMemsetAsync(d.prev_prec, 0, sizeof(int) * n_mol, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    ctx_release(ctx, b->dev_mem, b->dev_bytes);
    delete b;
    return fail(SG_ERR_HIP, std::string("batch upload: ") + hipGetErrorString(e));
  }
  b->e_cc.assign(n_mol, 0.0);
  *out = b;
  return SG_OK;
}

int sg_batch_destroy(sg_batch* b) {
  if (!b) return SG_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  (void)hipStreamSynchronize(b->ctx->side);
  drop_graphs(b);
  if (b->dev_mem) ctx_release(b->ctx, b->dev_mem, b->dev_bytes);
  if (b->diis_mem) (void)hipFree(b->diis_mem);
  if (b->hist_mem) (void)hipFree(b->hist_mem);
  delete b;
  return SG_OK;
}

int sg_batch_molecule_info(const sg_batch* b, int mol, int* n_atoms, int* n_ao, int* n_occ) {
  if (!b || mol < 0 || mol >= b->d.n_mol) return fail(SG_ERR_INVALID, "bad molecule index");
  if (n_atoms) *n_atoms = b->mols[mol].n_atoms;
  if (n_ao) *n_ao = b->mols[mol].n_ao;
  if (n_occ) *n_occ = b->mols[mol].n_occ;
  return SG_OK;
}

int sg_two_center(sg_batch* b) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  {
    StageTimer t(b->ctx, ST_TWO_CENTER);
    launch_pair_integrals(b->d, b->ctx->stream);
  }
  int r = check_launch("two-center integrals");
  if (r) return r;
  b->have_ints = true;
  b->have_fp32 = true;  // the integral kernels write the FP32 copy as well
  return SG_OK;
}

int sg_hcore(sg_batch* b) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  if (!b->have_ints) return fail(SG_ERR_ORDER, "sg_hcore needs sg_two_center first");
  HIP_TRY(hipSetDevice(b->ctx->device));
  {
    StageTimer t(b->ctx, ST_HCORE);
    launch_hcore(b->d, b->ctx->stream);
  }
  int r = check_launch("hcore");
  if (r) return r;
  b->have_hcore = true;
  return SG_OK;
}

int sg_core_core(sg_batch* b, double* e_cc) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  if (!b->have_ints) return fail(SG_ERR_ORDER, "sg_core_core needs sg_two_center first");
  HIP_TRY(hipSetDevice(b->ctx->device));
  {
    StageTimer t(b->ctx, ST_HCORE);
    launch_core_core(b->d, b->ctx->stream);
  }
  int r = check_launch("core-core");
  if (r) return r;
  HIP_TRY(hipMemcpyAsync(b->e_cc.data(), b->d.e_cc, sizeof(double) * b->d.n_mol,
                         hipMemcpyDeviceToHost, b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  b->have_ecc = true;
  if (e_cc) memcpy(e_cc, b->e_cc.data(), sizeof(double) * b->d.n_mol);
  return SG_OK;
}

int sg_guess_density(sg_batch* b) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  {
    StageTimer t(b->ctx, ST_GUESS);
    launch_guess_density(b->d, b->ctx->stream);
  }
  return check_launch("guess density");
}

int sg_fock(sg_batch* b) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  if (!b->have_hcore) return fail(SG_ERR_ORDER, "sg_fock needs sg_hcore first");
  HIP_TRY(hipSetDevice(b->ctx->device));
  {
    StageTimer t(b->ctx, ST_FOCK);
    launch_fock(b->d, b->ctx->stream, true);
  }
  return check_launch("fock");
}

int sg_eigh(sg_batch* b) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipMemsetAsync(b->d.status, 0, sizeof(int) * b->d.n_mol, b->ctx->stream));
  {
    timed_eigh(b->ctx, b->d, b->ctx->stream, true);
  }
  return check_launch("eigh");
}

int sg_eigh_noalloc(sg_context* c, int n, const double* a_dev, double* eigval_dev,
                    double* eigvec_dev, void* stream) {
  if (!c || !a_dev || !eigval_dev || !eigvec_dev)
    return fail(SG_ERR_INVALID, "sg_eigh_noalloc: null argument");
  if (n < 1 || n > kNoAllocMaxN)
    return fail(SG_ERR_INVALID, "sg_eigh_noalloc: n = " + std::to_string(n) + " outside 1.." +
                                    std::to_string(kNoAllocMaxN) +
                                    " (the fixed workspace's limit; sg_eigh takes larger ones)");
  if (c->device >= 64 || !g_scratch_ptr[c->device])
    return fail(SG_ERR_INVALID, "sg_eigh_noalloc: no workspace address for this device");
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  char* base = (char*)g_scratch_ptr[c->device];
  BatchDev d;
  d.n_mol = 1;
  d.n_mat = (int64_t)n * n;
  d.n_vec = n;
  d.max_n = n;
  d.max_n_small = n;
  set_tri_buckets(d, &n, 1);
  for (int k = 0; k < kTriBuckets; ++k) d.tri_variant[k] = c->tri_variant[k];
  d.mols = (MolDesc*)base;
  int* order = (int*)(base + 256);
  d.order = order;
  d.status = (int*)(base + 512);
  d.tnorm = (double*)(base + 768);
  double* v = (double*)(base + kNoAllocHead);
  d.dT = v;
  d.eT = v + kNoAllocMaxN;
  d.tau = v + 2 * kNoAllocMaxN;
  d.W = v + 3 * kNoAllocMaxN;
  d.F = const_cast<double*>(a_dev);
  d.eig = eigval_dev;
  d.C = eigvec_dev;
  static_assert(sizeof(MolDesc) <= 256, "descriptor room");
  k_noalloc_setup<<<1, 64, 0, s>>>(d.mols, order, d.status, n);
  int r = check_launch("sg_eigh_noalloc setup");
  if (r) return r;
  launch_eigh(d, s, true);
  return check_launch("sg_eigh_noalloc");
}

int sg_density(sg_batch* b) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  {
    StageTimer t(b->ctx, ST_DENSITY);
    launch_density(b->d, b->ctx->stream, true);
  }
  return check_launch("density");
}

int sg_energy(sg_batch* b, double* e_el) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  {
    StageTimer t(b->ctx, ST_ENERGY);
    launch_energy(b->d, b->ctx->stream, b->d.e_new, true);
  }
  int r = check_launch("energy");
  if (r) return r;
  if (e_el) {
    HIP_TRY(hipMemcpyAsync(e_el, b->d.e_new, sizeof(double) * b->d.n_mol, hipMemcpyDeviceToHost,
                           b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  }
  return SG_OK;
}

int sg_scf_run(sg_batch* b, double threshold, int max_steps, sg_scf_result* res) {
  return sg_scf_run_policy(b, threshold, max_steps, SG_POLICY_FP64, 1e-3, res);
}

int sg_scf_set_damping(sg_batch* b, double beta) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  if (!(beta >= 0.0 && beta < 1.0)) return fail(SG_ERR_INVALID, "damping must lie in [0, 1)");
  b->d.damping = beta;
  return SG_OK;
}

int sg_scf_record_energies(sg_batch* b, int cap) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  if (cap < 0 || cap > 100000) return fail(SG_ERR_INVALID, "sg_scf_record_energies: cap out of range");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  // the captured iteration holds the old record pointer
  drop_graphs(b);
  if (b->hist_mem) (void)hipFree(b->hist_mem);
  b->hist_mem = nullptr;
  b->d.e_hist = nullptr;
  b->d.e_hist_cap = 0;
  if (cap == 0) return SG_OK;
  const size_t bytes = sizeof(double) * (size_t)cap * (size_t)b->d.n_mol;
  hipError_t e = hipMalloc(&b->hist_mem, bytes);
  if (e != hipSuccess) {
    b->hist_mem = nullptr;
    return fail(SG_ERR_HIP, std::string("hipMalloc for the energy record: ") + hipGetErrorString(e));
  }
  HIP_TRY(hipMemset(b->hist_mem, 0, bytes));
  b->d.e_hist = (double*)b->hist_mem;
  b->d.e_hist_cap = cap;
  return SG_OK;
}

int sg_scf_energy_history(const sg_batch* b, int mol, double* out, int* count) {
  if (!b || !out || !count) return fail(SG_ERR_INVALID, "null argument");
  if (mol < 0 || mol >= b->d.n_mol) return fail(SG_ERR_INVALID, "bad molecule index");
  if (!b->d.e_hist) return fail(SG_ERR_ORDER, "sg_scf_energy_history: no record (sg_scf_record_energies)");
  HIP_TRY(hipSetDevice(b->ctx->device));
  hipStream_t s = b->ctx->stream;
  const int cap = b->d.e_hist_cap;
  int steps = 0;
  HIP_TRY(hipMemcpyAsync(out, b->d.e_hist + (size_t)mol * cap, sizeof(double) * cap,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&steps, b->d.steps + mol, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *count = std::min(steps + 1, cap);
  return SG_OK;
}

int sg_scf_set_diis(sg_batch* b, int depth) {
  if (!b) return fail(SG_ERR_INVALID, "null batch");
  if (depth != 0 && (depth < 2 || depth > kDiisMax))
    return fail(SG_ERR_INVALID, "DIIS depth must be 0 or lie in [2, " + std::to_string(kDiisMax) + "]");
  if (depth == b->d.diis_m) return SG_OK;
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  // the captured iteration holds the old depth and buffers
  drop_graphs(b);
  if (b->diis_mem) (void)hipFree(b->diis_mem);
  b->diis_mem = nullptr;
  BatchDev& d = b->d;
  d.diis_m = 0;
  d.diis_F = d.diis_E = d.diis_B = d.diis_part = d.diis_c = nullptr;
  d.diis_k = d.diis_p = nullptr;
  if (depth == 0) return SG_OK;
  const size_t n_mol = (size_t)d.n_mol;
  const size_t hist = sizeof(double) * (size_t)depth * (size_t)d.n_mat;
  const size_t o_F = 0, o_E = o_F + hist, o_B = o_E + hist,
               o_part = o_B + sizeof(double) * n_mol * kDiisMax * kDiisMax,
               o_c = o_part + sizeof(double) * (size_t)diis_part_doubles(d.n_mol),
               o_k = o_c + sizeof(double) * n_mol * kDiisMax, o_p = o_k + sizeof(int) * n_mol,
               end = o_p + sizeof(int) * n_mol;
  hipError_t e = hipMalloc(&b->diis_mem, end ? end : 256);
  if (e != hipSuccess) {
    b->diis_mem = nullptr;
    return fail(SG_ERR_HIP, std::string("hipMalloc(") + std::to_string(end) + ") for DIIS: " +
                                hipGetErrorString(e));
  }
  char* base = (char*)b->diis_mem;
  d.diis_F = (double*)(base + o_F);
  d.diis_E = (double*)(base + o_E);
  d.diis_B = (double*)(base + o_B);
  d.diis_part = (double*)(base + o_part);
  d.diis_c = (double*)(base + o_c);
  d.diis_k = (int*)(base + o_k);
  d.diis_p = (int*)(base + o_p);
  d.diis_stride = d.n_mat;
  d.diis_m = depth;
  return SG_OK;
}

int sg_scf_fp32_steps(const sg_batch* b, int* steps32) {
  if (!b || !steps32) return fail(SG_ERR_INVALID, "null batch or output");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipMemcpyAsync(steps32, b->d.steps32, sizeof(int) * b->d.n_mol, hipMemcpyDeviceToHost,
                         b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return SG_OK;
}

int sg_scf_run_policy(sg_batch* b, double threshold, int max_steps, int policy, double switch_ev,
                      sg_scf_result* res) {
  if (!b || max_steps < 0 || !(threshold >= 0.0) || !valid_policy(policy) || !(switch_ev >= 0.0))
    return fail(SG_ERR_INVALID, "sg_scf_run: bad arguments");
  int r;
  if (!b->have_ints && (r = sg_two_center(b))) return r;
  if (!b->have_hcore && (r = sg_hcore(b))) return r;
  if (!b->have_ecc && (r = sg_core_core(b, nullptr))) return r;
  sg_context* c = b->ctx;
  hipStream_t s = c->stream;
  HIP_TRY(hipSetDevice(c->device));
  if (policy != SG_POLICY_FP64 && !b->have_fp32) {  // integrals set by the stage API only
    StageTimer t(c, ST_TWO_CENTER);
    launch_blocks_to_fp32(b->d, s);
    b->have_fp32 = true;
  }
  // InitialGuess + E0 (scfcalculation.cpp:66-72, 148-165), in FP64
  {
    StageTimer t(c, ST_GUESS);
    launch_guess_density(b->d, s);
  }
  {
    StageTimer t(c, ST_FOCK);
    launch_fock(b->d, s, true);
  }
  {
    StageTimer t(c, ST_ENERGY);
    launch_energy_fock(b->d, s, b->d.e_new, true);
  }
  {
    StageTimer t(c, ST_CONTROL);
    launch_set_prec(b->d, s, reduced_prec(policy));
    launch_scf_init(b->d, s, max_steps, policy);
    launch_diis_reset(b->d, s);
  }
  if ((r = check_launch("scf init"))) return r;
  const int fp32_cap = 40;
  b->d.mixed = policy == SG_POLICY_FP64 ? 0 : 1;
  b->d.eig32 = policy == SG_POLICY_MIXED_EIG32 ? 1 : 0;
  // The molecules iterate as one group on the context's stream, or as two
  // (split_views) on it and on the side stream, each with its own iteration
  // and active count, so that neither group waits for the other.
  BatchDev views[2];
  hipStream_t streams[2] = {s, c->side};
  int n_groups = 1;
  if (can_split(c, b->d)) {
    split_views(b->d, views[0], views[1]);
    n_groups = 2;
  } else {
    views[0] = b->d;
  }
  // Without stage timing, each group's iteration is captured once as a HIP
  // graph and replayed; its kernel arguments depend on the batch and on these
  // settings.  Batches with molecules on the multi-workgroup eigensolver
  // (thousands of launches per iteration) launch directly.
  const bool use_graph = c->graphs && !c->timing && b->d.n_large == 0;
  if (use_graph && (!b->iter_exec || b->g_threshold != threshold || b->g_max_steps != max_steps ||
                    b->g_switch != switch_ev || b->g_mixed != policy ||
                    b->g_damping != b->d.damping)) {
    drop_graphs(b);
    for (int g = 0; g < n_groups; ++g)
      if ((r = capture_iteration(c, views[g], streams[g], threshold, max_steps, switch_ev,
                                 fp32_cap, g ? &b->iter_exec2 : &b->iter_exec))) {
        drop_graphs(b);
        return r;
      }
    b->g_threshold = threshold;
    b->g_max_steps = max_steps;
    b->g_switch = switch_ev;
    b->g_mixed = policy;
    b->g_damping = b->d.damping;
  }
  if (n_groups == 2) {  // the side stream starts after the set-up above
    HIP_TRY(hipEventRecord(c->ev_fork, s));
    HIP_TRY(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  }
  // Each group keeps up to two iterations in flight on its stream: the host
  // reads a group's active count (a copy to pinned memory, then an event) one
  // iteration behind its launches, so the device does not idle while the host
  // reads.  The iteration launched after a group's count reached zero finds no
  // active molecule and changes nothing.  Multi-workgroup batches keep one.
  const int depth = b->d.n_large ? 1 : 2;
  int launched[2] = {0, 0}, read[2] = {0, 0};
  const bool none = max_steps == 0 || b->d.n_mol == 0;
  bool done[2] = {none, none};
  std::deque<int> fifo;  // the group of each iteration in flight, oldest first
  auto launch = [&](int g) -> int {
    const int slot = 2 * g + (launched[g] & 1);
    if (use_graph)
      HIP_TRY(hipGraphLaunch(g ? b->iter_exec2 : b->iter_exec, streams[g]));
    else
      scf_iteration(c, views[g], streams[g], threshold, max_steps, switch_ev, fp32_cap);
    int rc;
    if ((rc = check_launch("scf step"))) return rc;
    HIP_TRY(hipMemcpyAsync(c->host_counts + slot, views[g].n_active, sizeof(int),
                           hipMemcpyDeviceToHost, streams[g]));
    HIP_TRY(hipEventRecord(c->ev_count[slot], streams[g]));
    ++launched[g];
    fifo.push_back(g);
    return SG_OK;
  };
  for (int g = 0; g < n_groups; ++g)
    for (int k = 0; k < depth && !done[g] && launched[g] < max_steps; ++k)
      if ((r = launch(g))) return r;
  while (!fifo.empty()) {
    const int g = fifo.front();
    fifo.pop_front();
    const int slot = 2 * g + (read[g] & 1);
    HIP_TRY(hipEventSynchronize(c->ev_count[slot]));
    ++read[g];
    if (done[g]) continue;
    if (c->host_counts[slot] == 0) done[g] = true;
    if (!done[g] && launched[g] < max_steps && (r = launch(g))) return r;
  }
  if (n_groups == 2) {
    HIP_TRY(hipEventRecord(c->ev_join, c->side));
    HIP_TRY(hipStreamWaitEvent(s, c->ev_join, 0));
  }
  b->have_scf = true;
  if (res) {
    const int n = b->d.n_mol;
    std::vector<double> e(n);
    std::vector<int> st(n), cv(n), ss(n);
    HIP_TRY(hipMemcpyAsync(e.data(), b->d.e_el, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(st.data(), b->d.steps, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(cv.data(), b->d.converged, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ss.data(), b->d.status, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int m = 0; m < n; ++m) {
      res[m].e_el = e[m];
      res[m].e_cc = b->e_cc[m];
      res[m].e_tot = e[m] + b->e_cc[m];
      res[m].steps = st[m];
      res[m].converged = cv[m];
      res[m].status = ss[m];
    }
  }
  return SG_OK;
}

long sg_pair_block_count(const sg_batch* b, int mol) {
  if (!b || mol < 0 || mol >= b->d.n_mol) return -1;
  return (long)b->blk_count[mol];
}

int sg_get_pair_blocks(const sg_batch* b, int mol, double* out) {
  if (!b || !out || mol < 0 || mol >= b->d.n_mol) return fail(SG_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipMemcpyAsync(out, b->d.blocks + b->mols[mol].blk0, sizeof(double) * b->blk_count[mol],
                         hipMemcpyDeviceToHost, b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return SG_OK;
}

static double* matrix_ptr(const sg_batch* b, int which) {
  return which == SG_HCORE ? b->d.H : which == SG_FOCK ? b->d.F : which == SG_DENSITY ? b->d.P : nullptr;
}

int sg_get_packed(const sg_batch* b, int mol, int which, double* out) {
  if (!b || !out || mol < 0 || mol >= b->d.n_mol || !matrix_ptr(b, which))
    return fail(SG_ERR_INVALID, "bad arguments");
  const MolDesc& M = b->mols[mol];
  const int n = M.n_ao;
  std::vector<double> full((size_t)n * n);
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipMemcpyAsync(full.data(), matrix_ptr(b, which) + M.mat0, sizeof(double) * full.size(),
                         hipMemcpyDeviceToHost, b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) out[(size_t)i * (i + 1) / 2 + j] = full[(size_t)i * n + j];
  return SG_OK;
}

int sg_set_packed(sg_batch* b, int mol, int which, const double* in) {
  if (!b || !in || mol < 0 || mol >= b->d.n_mol || !matrix_ptr(b, which))
    return fail(SG_ERR_INVALID, "bad arguments");
  const MolDesc& M = b->mols[mol];
  const int n = M.n_ao;
  std::vector<double> full((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const double v = in[(size_t)i * (i + 1) / 2 + j];
      full[(size_t)i * n + j] = v;
      full[(size_t)j * n + i] = v;
    }
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipMemcpyAsync(matrix_ptr(b, which) + M.mat0, full.data(), sizeof(double) * full.size(),
                         hipMemcpyHostToDevice, b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  if (which == SG_HCORE) b->have_hcore = true;
  return SG_OK;
}

int sg_get_eigen(const sg_batch* b, int mol, double* eigval, double* eigvec) {
  if (!b || mol < 0 || mol >= b->d.n_mol) return fail(SG_ERR_INVALID, "bad arguments");
  const MolDesc& M = b->mols[mol];
  const size_t n = M.n_ao;
  HIP_TRY(hipSetDevice(b->ctx->device));
  if (eigval)
    HIP_TRY(hipMemcpyAsync(eigval, b->d.eig + M.vec0, sizeof(double) * n, hipMemcpyDeviceToHost,
                           b->ctx->stream));
  if (eigvec)
    HIP_TRY(hipMemcpyAsync(eigvec, b->d.C + M.mat0, sizeof(double) * n * n, hipMemcpyDeviceToHost,
                           b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return SG_OK;
}

int sg_set_pair_blocks(sg_batch* b, int mol, const double* in) {
  if (!b || !in || mol < 0 || mol >= b->d.n_mol) return fail(SG_ERR_INVALID, "bad arguments");
  const int64_t n = b->blk_count[mol];
  HIP_TRY(hipSetDevice(b->ctx->device));
  std::vector<float> f(n);
  for (int64_t i = 0; i < n; ++i) f[i] = (float)in[i];
  HIP_TRY(hipMemcpyAsync(b->d.blocks + b->mols[mol].blk0, in, sizeof(double) * n,
                         hipMemcpyHostToDevice, b->ctx->stream));
  HIP_TRY(hipMemcpyAsync(b->d.blocks32 + b->mols[mol].blk0, f.data(), sizeof(float) * n,
                         hipMemcpyHostToDevice, b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  b->have_ints = true;
  return SG_OK;
}

int sg_heat_of_formation(int n_atoms, const int* Z, double e_tot_ev, double* dhf_kcal) {
  if (n_atoms <= 0 || !Z || !dhf_kcal) return fail(SG_ERR_INVALID, "bad arguments");
  // MNDO atomic tables (MNDO_parameters.cpp:58-59, 80-81, 100-101, 120-121):
  // experimental heat of formation of the atom (kcal/mol) and the MNDO
  // energy of the isolated atom (eV); 1 eV = 23.060529 kcal/mol
  double e_isol = 0.0, h_atoms = 0.0;
  for (int i = 0; i < n_atoms; ++i) {
    double heat, eisol;
    switch (Z[i]) {
      case 1: heat = 52.102; eisol = -11.9062760; break;
      case 6: heat = 170.89; eisol = -120.500606; break;
      case 7: heat = 113.0; eisol = -202.566201; break;
      case 8: heat = 59.559; eisol = -317.868506; break;
      default:
        return fail(SG_ERR_UNSUPPORTED, "element Z=" + std::to_string(Z[i]) +
                                            " has no MNDO parameters here");
    }
    e_isol += eisol;
    h_atoms += heat;
  }
  *dhf_kcal = (e_tot_ev - e_isol) * 23.060529 + h_atoms;
  return SG_OK;
}

int sg_gradient_fd(sg_context* ctx, int n_atoms, const int* Z, const double* xyz_ang,
                   double step_ang, double threshold_ev, int diis_depth, double* e_tot,
                   double* grad, int* all_converged) {
  if (!ctx || n_atoms <= 0 || !Z || !xyz_ang || !grad || !(step_ang > 0.0) ||
      !(threshold_ev >= 0.0))
    return fail(SG_ERR_INVALID, "sg_gradient_fd: bad arguments");
  // molecule 0: the geometry; 1 + 2 (3 a + k) and 2 + 2 (3 a + k): atom a
  // moved by +step and -step along axis k
  const int n_mol = 1 + 6 * n_atoms;
  std::vector<int> offs(n_mol + 1), zz((size_t)n_mol * n_atoms);
  std::vector<double> xyz((size_t)n_mol * n_atoms * 3);
  for (int m = 0; m < n_mol; ++m) {
    offs[m] = m * n_atoms;
    std::copy(Z, Z + n_atoms, zz.begin() + (size_t)m * n_atoms);
    std::copy(xyz_ang, xyz_ang + 3 * n_atoms, xyz.begin() + (size_t)m * n_atoms * 3);
    if (m) {
      const int q = (m - 1) / 2, a = q / 3, k = q % 3;
      xyz[((size_t)m * n_atoms + a) * 3 + k] += (m - 1) % 2 == 0 ? step_ang : -step_ang;
    }
  }
  offs[n_mol] = n_mol * n_atoms;
  sg_batch* b = nullptr;
  int r = sg_batch_create(ctx, n_mol, offs.data(), zz.data(), xyz.data(), &b);
  if (r) return r;
  std::vector<sg_scf_result> res(n_mol);
  r = diis_depth ? sg_scf_set_diis(b, diis_depth) : SG_OK;
  if (!r) r = sg_scf_run_policy(b, threshold_ev, 200, SG_POLICY_FP64, 1e-3, res.data());
  sg_batch_destroy(b);
  if (r) return r;
  int conv = 1;
  for (const sg_scf_result& x : res) conv = conv && x.converged && x.status == 0;
  for (int q = 0; q < 3 * n_atoms; ++q)
    grad[q] = (res[1 + 2 * q].e_tot - res[2 + 2 * q].e_tot) / (2.0 * step_ang);
  if (e_tot) *e_tot = res[0].e_tot;
  if (all_converged) *all_converged = conv;
  return SG_OK;
}

int sg_gradient(sg_batch* b, double* grad) {
  if (!b || !grad) return fail(SG_ERR_INVALID, "sg_gradient: null batch or output");
  if (!b->have_scf) return fail(SG_ERR_ORDER, "sg_gradient needs an SCF first (sg_scf_run)");
  HIP_TRY(hipSetDevice(b->ctx->device));
  hipStream_t s = b->ctx->stream;
  // the output passes through the per-atom Fock weights (10 doubles per atom,
  // rebuilt by every Fock build)
  launch_gradient(b->d, s, b->d.wdiag);
  int r = check_launch("gradient");
  if (r) return r;
  HIP_TRY(hipMemcpyAsync(grad, b->d.wdiag, sizeof(double) * 3 * (size_t)b->d.n_atoms,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return SG_OK;
}

int sg_get_status(const sg_batch* b, int* status) {
  if (!b || !status) return fail(SG_ERR_INVALID, "bad arguments");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipMemcpyAsync(status, b->d.status, sizeof(int) * b->d.n_mol, hipMemcpyDeviceToHost,
                         b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return SG_OK;
}

int sg_set_eigen(sg_batch* b, int mol, const double* eigval, const double* eigvec) {
  if (!b || mol < 0 || mol >= b->d.n_mol) return fail(SG_ERR_INVALID, "bad arguments");
  const MolDesc& M = b->mols[mol];
  const size_t n = M.n_ao;
  HIP_TRY(hipSetDevice(b->ctx->device));
  if (eigval)
    HIP_TRY(hipMemcpyAsync(b->d.eig + M.vec0, eigval, sizeof(double) * n, hipMemcpyHostToDevice,
                           b->ctx->stream));
  if (eigvec)
    HIP_TRY(hipMemcpyAsync(b->d.C + M.mat0, eigvec, sizeof(double) * n * n, hipMemcpyHostToDevice,
                           b->ctx->stream));
  HIP_TRY(hipStreamSynchronize(b->ctx->stream));
  return SG_OK;
}

}  // extern "C"
