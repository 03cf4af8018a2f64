// sg_server.cpp -- continuous-batching SCF server (C ABI in include/semigpu.h).
//
// Requests are routed by basis size N to one of the server's size classes
// ("lanes").  The server owns fixed-capacity device pools carved into the
// same arrays the batch kernels use (N^2 matrix space for H, F, P, C and the
// eigensolver workspace; pair-block space, FP64 and FP32; length-N vectors),
// shared by all lanes with first-fit allocation.  A lane owns its molecule
// slots, atom descriptors, pair-list scratch, HIP stream and pinned host
// mirrors.  Lanes run independently: a small molecule's iterations are paced
// by the molecules of its own class, not by the largest one in flight, and
// the lanes' kernels overlap on the GPU.
//
// sg_server_step() visits every lane.  A lane whose last iteration has not
// finished (event query) is skipped.  Otherwise it
//   0. retires the molecules that converged, hit the step cap or failed in that
//      iteration (from the status read back into pinned memory), frees their
//      ranges and slot and records the result with its timestamps;
//   1. admits queued requests in arrival order while their footprint fits the
//      free ranges (first fit) and a slot is free;
//   2. runs integrals, Hcore, core-core, guess density, guess Fock and E_0 for
//      the newly admitted molecules only (per-molecule mask);
//   3. launches one SCF iteration for every molecule in flight (eigensolver,
//      density, Fock, energy, convergence bookkeeping) and the asynchronous
//      readback of their status, then records the lane's event.
// The call sleeps ~100 us when no lane moved, so polling loops do not spin.
// State of molecules in flight stays in HBM between steps.  Slot index
// max_slots of a lane is a permanent dead slot (never active, never masked):
// atom descriptors of freed atom ranges point to it, so kernels that walk the
// atom array skip them.  Every kernel works per molecule with a fixed
// summation order, so a molecule's result does not depend on what else is in
// flight, in its lane or any other.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "sg_host.h"

using namespace sg;

namespace {

// First-fit allocator over [0, capacity) with coalescing on release.
class RangeAlloc {
 public:
  void reset(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  bool alloc(int64_t n, int64_t* off) {
    if (n <= 0) {
      *off = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const int64_t rest = it->second - n, start = it->first + n;
        free_.erase(it);
        if (rest) free_[start] = rest;
        return true;
      }
    }
    return false;
  }
  void release(int64_t off, int64_t n) {
    if (n <= 0) return;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free_[off] = n;
  }

 private:
  std::map<int64_t, int64_t> free_;
};

// Page-locked host array (hipHostMalloc), the source or target of the lane's
// asynchronous copies.
template <typename T>
struct Pinned {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    n = count;
    return hipHostMalloc((void**)&p, sizeof(T) * (count ? count : 1), hipHostMallocDefault);
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  T& operator[](size_t i) { return p[i]; }
};

struct Request {
  long long id = 0;
  std::vector<int> Z;
  std::vector<double> xyz;
  double t_submit = 0.0;
  int n_ao = 0, n_occ = 0;
  int64_t n_blk = 0;
};

struct Slot {
  bool used = false;
  Request req;
  double t_admit = 0.0;
  int64_t mat0 = 0, blk0 = 0, vec0 = 0, atom0 = 0;
};

int64_t pair_block_doubles(const std::vector<int>& Z) {
  int64_t s = 0;
  for (size_t i = 0; i < Z.size(); ++i)
    for (size_t j = 0; j <= i; ++j) s += block_size(Z[i], Z[j]);
  return s;
}

// One size class: molecules with n_lo <= N < n_hi.  Matrix, pair-block and
// vector space come from the server's shared pools; slots, atoms and the
// per-molecule scalars are the lane's own.
struct Lane {
  int n_lo = 0, n_hi = 0;
  int max_slots = 0, max_atoms = 0;
  int64_t pair_cap = 0;  // scratch entries for the pair lists set up per step
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  bool own_stream = false;  // created here (else one of the context's class streams)
  void* dev = nullptr;
  BatchDev d;
  int* mask_new = nullptr;  // device, per slot (+ dead slot)
  int2* pair_scratch = nullptr;
  int* order_dev = nullptr;  // slots by decreasing N (device)
  bool order_dirty = true;
  // pinned mirrors: descriptors, mask, order, pair-list staging, readback
  Pinned<MolDesc> mols;
  Pinned<AtomDesc> atoms;
  Pinned<AtomDesc> partners;  // the Fock build's partner order per molecule (descriptor copies)
  Pinned<int> mask, order;
  std::vector<int> tri_n;  // N of the slots in launch order (reduction buckets)
  std::vector<int> slot_used;  // slots that hold a molecule (launch shapes)
  Pinned<int2> pairs;
  Pinned<int> rb_act, rb_steps, rb_conv, rb_stat;
  Pinned<double> rb_e, rb_ecc;
  std::set<int> free_slots;
  std::vector<Slot> slots;
  std::deque<Request> queue;
  RangeAlloc atom;
  bool busy = false;        // an iteration and its readback are in flight
  int hw = 0;               // slots covered by that iteration
  void* diis_dev = nullptr;  // DIIS B matrices, partial sums, coefficients, counters per slot
  std::vector<int> fresh;   // slots it set up
};

}  // namespace

struct sg_server {
  sg_context* ctx = nullptr;
  double threshold = 5e-3, switch_ev = 1e-3;
  int max_steps = 100, policy = SG_POLICY_FP64;
  // shared pools: N^2 matrices (H, F, P, C, 5 N^2 eigensolver workspace),
  // pair blocks (FP64 + FP32), length-N vectors; first fit over each
  void* pool = nullptr;
  void* diis_pool = nullptr;  // DIIS histories over the matrix pool (sg_server_set_diis)
  BatchDev shared;
  int64_t mat_cap = 0, blk_cap = 0, vec_cap = 0;
  RangeAlloc mat, blk, vec;
  std::vector<Lane*> lanes;
  std::vector<sg_served> done;
  std::chrono::steady_clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  ~sg_server() {
    for (Lane* L : lanes) {
      if (L->stream) (void)hipStreamSynchronize(L->stream);
      if (L->dev) (void)hipFree(L->dev);
      if (L->diis_dev) (void)hipFree(L->diis_dev);
      if (L->ev) (void)hipEventDestroy(L->ev);
      if (L->stream && L->own_stream) (void)hipStreamDestroy(L->stream);
      delete L;
    }
    if (pool) (void)hipFree(pool);
    if (diis_pool) (void)hipFree(diis_pool);
  }
};

namespace {

// Per-lane device memory: slot descriptors and scalars, atoms, pair scratch.
size_t lane_bytes(int max_slots, int max_atoms, int64_t pair_cap) {
  const size_t ns = max_slots + 1;
  return sizeof(MolDesc) * ns + (sizeof(AtomDesc) + 12) * max_atoms + sizeof(int2) * pair_cap +
         ns * (5 * 8 + 9 * 4) + 256 + 16 * 256;
}

int lane_create(const BatchDev& shared, Lane* L, int max_slots, int max_atoms, int64_t pair_cap,
                hipStream_t stream) {
  L->max_slots = max_slots;
  L->max_atoms = max_atoms;
  L->pair_cap = pair_cap;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t ns = max_slots + 1;  // + dead slot
  const size_t o_mols = take(sizeof(MolDesc) * ns), o_atoms = take(sizeof(AtomDesc) * max_atoms);
  const size_t o_eatom = take(8 * (size_t)max_atoms),
               o_partners = take(sizeof(AtomDesc) * (size_t)max_atoms);
  const size_t o_pairs = take(sizeof(int2) * L->pair_cap);
  const size_t o_tn = take(8 * ns), o_ecc = take(8 * ns), o_eel = take(8 * ns), o_enew = take(8 * ns);
  const size_t o_steps = take(4 * ns), o_act = take(4 * ns), o_conv = take(4 * ns),
               o_stat = take(4 * ns), o_nact = take(256), o_prec = take(4 * ns),
               o_pprec = take(4 * ns), o_mask = take(4 * ns), o_order = take(4 * ns),
               o_dep = take(8 * ns), o_wdiag = take(80 * (size_t)max_atoms);
  if (stream) {
    L->stream = stream;
  } else {
    HIP_TRY(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
    L->own_stream = true;
  }
  HIP_TRY(hipEventCreateWithFlags(&L->ev, hipEventDisableTiming));
  hipError_t e = hipMalloc(&L->dev, off);
  if (e != hipSuccess) {
    L->dev = nullptr;
    return fail(SG_ERR_HIP, std::string("sg_server_create: hipMalloc(") + std::to_string(off) +
                                "): " + hipGetErrorString(e));
  }
  char* base = (char*)L->dev;
  BatchDev& d = L->d;
  d = shared;
  d.mols = (MolDesc*)(base + o_mols);
  d.atoms = (AtomDesc*)(base + o_atoms);
  d.e_atom = (double*)(base + o_eatom);
  d.partners = (AtomDesc*)(base + o_partners);
  L->pair_scratch = (int2*)(base + o_pairs);
  d.tnorm = (double*)(base + o_tn);
  d.e_cc = (double*)(base + o_ecc);
  d.e_el = (double*)(base + o_eel);
  d.e_new = (double*)(base + o_enew);
  d.steps = (int*)(base + o_steps);
  d.active = (int*)(base + o_act);
  d.converged = (int*)(base + o_conv);
  d.status = (int*)(base + o_stat);
  d.n_active = (int*)(base + o_nact);
  d.prec = (int*)(base + o_prec);
  d.prev_prec = (int*)(base + o_pprec);
  d.de_prev = (double*)(base + o_dep);
  d.wdiag = (double*)(base + o_wdiag);
  L->mask_new = (int*)(base + o_mask);
  L->order_dev = (int*)(base + o_order);
  HIP_TRY(L->mols.alloc(ns));
  HIP_TRY(L->atoms.alloc(max_atoms));
  HIP_TRY(L->partners.alloc(max_atoms));
  HIP_TRY(L->mask.alloc(ns));
  HIP_TRY(L->order.alloc(ns));
  L->tri_n.assign(ns, 0);
  HIP_TRY(L->pairs.alloc(L->pair_cap));
  for (Pinned<int>* p : {&L->rb_act, &L->rb_steps, &L->rb_conv, &L->rb_stat}) HIP_TRY(p->alloc(ns));
  HIP_TRY(L->rb_e.alloc(ns));
  HIP_TRY(L->rb_ecc.alloc(ns));
  hipStream_t s = L->stream;
  for (int* p : {d.steps, d.active, d.converged, d.status, d.prec, d.prev_prec, L->mask_new})
    HIP_TRY(hipMemsetAsync(p, 0, 4 * ns, s));
  for (size_t i = 0; i < ns; ++i) {
    L->mols[i] = MolDesc{};
    L->mask[i] = 0;
  }
  AtomDesc dead{};
  dead.mol = max_slots;
  for (int i = 0; i < max_atoms; ++i) L->atoms[i] = dead;
  HIP_TRY(hipMemcpyAsync(d.mols, L->mols.p, sizeof(MolDesc) * ns, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d.atoms, L->atoms.p, sizeof(AtomDesc) * max_atoms, hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  L->slots.assign(max_slots, Slot{});
  for (int i = 0; i < max_slots; ++i) L->free_slots.insert(i);
  L->atom.reset(max_atoms);
  return SG_OK;
}

// 0. results of the iteration that just finished
int lane_finish(sg_server* S, Lane* L) {
  hipStream_t s = L->stream;
  for (int sl : L->fresh) L->mask[sl] = 0;
  L->fresh.clear();
  const double tdone = S->now();
  for (int i = 0; i < L->hw; ++i) {
    Slot& sl = L->slots[i];
    if (!sl.used || L->rb_act[i]) continue;
    sg_served o;
    memset(&o, 0, sizeof(o));
    o.request_id = sl.req.id;
    o.e_el = L->rb_e[i];
    o.e_cc = L->rb_ecc[i];
    o.e_tot = L->rb_e[i] + L->rb_ecc[i];
    o.steps = L->rb_steps[i];
    o.converged = L->rb_conv[i];
    o.status = L->rb_stat[i];
    o.n_ao = sl.req.n_ao;
    o.t_submit = sl.req.t_submit;
    o.t_admit = sl.t_admit;
    o.t_done = tdone;
    S->done.push_back(o);
    const MolDesc& M = L->mols[i];
    S->mat.release(sl.mat0, (int64_t)M.n_ao * M.n_ao);
    S->blk.release(sl.blk0, sl.req.n_blk);
    S->vec.release(sl.vec0, M.n_ao);
    L->atom.release(sl.atom0, M.n_atoms);
    AtomDesc dead{};
    dead.mol = L->max_slots;
    for (int a = 0; a < M.n_atoms; ++a) L->atoms[M.atom0 + a] = dead;
    // the copy reads the pinned mirror when it runs: an admission in this
    // same step that reuses part of the range rewrites those entries first and
    // enqueues its own copy behind this one, so the device ends with the new
    // descriptors either way
    HIP_TRY(hipMemcpyAsync(L->d.atoms + M.atom0, L->atoms.p + M.atom0,
                           sizeof(AtomDesc) * M.n_atoms, hipMemcpyHostToDevice, s));
    L->free_slots.insert(i);
    sl = Slot{};
    L->order_dirty = true;
  }
  L->busy = false;
  return SG_OK;
}

// 1-3. admission, setup of the new molecules, one iteration for all
int lane_launch(sg_server* S, Lane* L) {
  sg_context* c = S->ctx;
  hipStream_t s = L->stream;
  std::vector<int2> pxx, phx, phh, pall, diag;
  int64_t pairs_budget = L->pair_cap;
  while (!L->queue.empty()) {
    Request& r = L->queue.front();
    const int na = (int)r.Z.size();
    // scratch entries: each atom pair twice (type list + full list), <= 10
    // Hcore tasks per heavy atom (4 AOs) and 1 per hydrogen, i.e. < 3 N
    const int64_t need = (int64_t)na * (na - 1) + 3 * (int64_t)r.n_ao;
    if (need > pairs_budget || L->free_slots.empty()) break;
    const int slot = *L->free_slots.begin();
    int64_t m0, b0, v0, a0;
    const int64_t n2 = (int64_t)r.n_ao * r.n_ao;
    if (!S->mat.alloc(n2, &m0)) break;
    if (!S->blk.alloc(r.n_blk, &b0)) {
      S->mat.release(m0, n2);
      break;
    }
    if (!S->vec.alloc(r.n_ao, &v0)) {
      S->mat.release(m0, n2);
      S->blk.release(b0, r.n_blk);
      break;
    }
    if (!L->atom.alloc(na, &a0)) {
      S->mat.release(m0, n2);
      S->blk.release(b0, r.n_blk);
      S->vec.release(v0, r.n_ao);
      break;
    }
    L->free_slots.erase(L->free_slots.begin());
    Slot& sl = L->slots[slot];
    sl.used = true;
    sl.t_admit = S->now();
    sl.mat0 = m0;
    sl.blk0 = b0;
    sl.vec0 = v0;
    sl.atom0 = a0;
    MolDesc& M = L->mols[slot];
    M.atom0 = (int)a0;
    M.n_atoms = na;
    M.n_ao = r.n_ao;
    M.n_occ = r.n_occ;
    M.mat0 = m0;
    M.blk0 = b0;
    M.vec0 = v0;
    int ao = 0, hv = 0;
    int64_t row = 0;
    for (int a = 0; a < na; ++a) {
      AtomDesc& A = L->atoms[a0 + a];
      A.zn = r.Z[a];
      A.x = r.xyz[3 * a + 0] * 1.8897259886;
      A.y = r.xyz[3 * a + 1] * 1.8897259886;
      A.z = r.xyz[3 * a + 2] * 1.8897259886;
      A.mol = slot;
      A.local = a;
      A.ao0 = ao;
      A.hv_before = hv;
      A.pad = 0;
      A.row0 = row;
      for (int j = 0; j <= a; ++j) row += block_size(A.zn, r.Z[j]);
      ao += n_ao_of(A.zn);
      if (A.zn > 1) ++hv;
      const int gi = (int)a0 + a;
      for (int j = 0; j < a; ++j) {
        int2 pr = make_int2(gi, (int)a0 + j);
        pall.push_back(pr);
        const int bs = block_size(A.zn, r.Z[j]);
        (bs == 100 ? pxx : bs == 10 ? phx : phh).push_back(pr);
      }
      for (int t = 0; t < (A.zn > 1 ? 10 : 1); ++t) diag.push_back(make_int2(gi, t));
    }
    // partner order of the Fock build: heavy atoms, then hydrogens
    int np = 0;
    for (int pass = 0; pass < 2; ++pass)
      for (int a = 0; a < na; ++a)
        if ((r.Z[a] > 1) == (pass == 0)) L->partners[a0 + np++] = L->atoms[a0 + a];
    sl.req = std::move(r);
    L->queue.pop_front();
    L->fresh.push_back(slot);
    pairs_budget -= need;
  }
  // high-water marks of the slot and atom ranges in use
  int hw = 0, hw_atoms = 0, max_n = 1;
  for (int i = 0; i < L->max_slots; ++i)
    if (L->slots[i].used) {
      const MolDesc& M = L->mols[i];
      hw = i + 1;
      hw_atoms = std::max(hw_atoms, M.atom0 + M.n_atoms);
      max_n = std::max(max_n, M.n_ao);
    }
  L->hw = hw;
  if (hw == 0) return SG_OK;
  BatchDev it = L->d;
  for (int k = 0; k < kTriBuckets; ++k) it.tri_variant[k] = S->ctx->tri_variant[k];
  it.n_mol = hw;
  it.n_atoms = hw_atoms;
  it.max_n = max_n;
  {
    std::vector<int>& used = L->slot_used;
    used.assign(hw, 0);
    for (int i = 0; i < hw; ++i) used[i] = L->slots[i].used ? 1 : 0;
    set_launch_shapes(it, L->mols.p, hw, used.data());
  }
  if (!L->fresh.empty()) L->order_dirty = true;
  if (L->order_dirty) {
    // slots in flight by decreasing N, then the free slots below the high-water mark
    for (int i = 0; i < hw; ++i) L->order[i] = i;
    std::stable_sort(L->order.p, L->order.p + hw, [&](int x, int y) {
      const int nx = L->slots[x].used ? L->mols[x].n_ao : -1;
      const int ny = L->slots[y].used ? L->mols[y].n_ao : -1;
      return nx > ny;
    });
    HIP_TRY(hipMemcpyAsync(L->order_dev, L->order.p, sizeof(int) * hw, hipMemcpyHostToDevice, s));
    L->order_dirty = false;
  }
  it.order = L->order_dev;
  {
    // reduction buckets from the slots' N in launch order (free slots: 0)
    int* nb = L->tri_n.data();
    for (int i = 0; i < hw; ++i) {
      const int sl = L->order[i];
      nb[i] = L->slots[sl].used ? L->mols[sl].n_ao : 0;
    }
    set_tri_buckets(it, nb, hw);
  }
  if (!L->fresh.empty()) {
    for (int sl : L->fresh) L->mask[sl] = 1;
    HIP_TRY(hipMemcpyAsync(L->d.mols, L->mols.p, sizeof(MolDesc) * hw, hipMemcpyHostToDevice, s));
    for (int sl : L->fresh) {
      const MolDesc& M = L->mols[sl];
      HIP_TRY(hipMemcpyAsync(L->d.atoms + M.atom0, L->atoms.p + M.atom0,
                             sizeof(AtomDesc) * M.n_atoms, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(L->d.partners + M.atom0, L->partners.p + M.atom0,
                             sizeof(AtomDesc) * M.n_atoms, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipMemcpyAsync(L->mask_new, L->mask.p, sizeof(int) * (L->max_slots + 1),
                           hipMemcpyHostToDevice, s));
    // pair lists and Hcore tasks of the new molecules, staged in pinned memory
    // and packed into the device scratch
    BatchDev nb = it;
    size_t used = 0;
    auto put = [&](const std::vector<int2>& v, int2*& dst, int& cnt) {
      dst = L->pair_scratch + used;
      cnt = (int)v.size();
      std::copy(v.begin(), v.end(), L->pairs.p + used);
      used += v.size();
    };
    put(pxx, nb.pairs_xx, nb.n_xx);
    put(phx, nb.pairs_hx, nb.n_hx);
    put(phh, nb.pairs_hh, nb.n_hh);
    put(pall, nb.pairs_all, nb.n_pairs_all);
    put(diag, nb.diag_tasks, nb.n_diag_tasks);
    if (used)
      HIP_TRY(hipMemcpyAsync(L->pair_scratch, L->pairs.p, sizeof(int2) * used,
                             hipMemcpyHostToDevice, s));
    {
      StageTimer t(c, ST_TWO_CENTER, s);
      launch_pair_integrals(nb, s, L->mask_new);
    }
    {
      StageTimer t(c, ST_HCORE, s);
      launch_clear_hf(nb, s, L->mask_new);
      launch_hcore(nb, s);
      launch_core_core(nb, s, L->mask_new);
    }
    {
      StageTimer t(c, ST_GUESS, s);
      launch_guess_density(nb, s, L->mask_new);
    }
    {
      StageTimer t(c, ST_FOCK, s);
      launch_fock(nb, s, true, L->mask_new);
    }
    {
      StageTimer t(c, ST_ENERGY, s);
      launch_energy_fock(nb, s, nb.e_new, true, L->mask_new);
    }
    {
      StageTimer t(c, ST_CONTROL, s);
      launch_set_prec(nb, s, reduced_prec(S->policy), L->mask_new);
      launch_scf_init(nb, s, S->max_steps, S->policy, L->mask_new);
      launch_diis_reset(nb, s, L->mask_new);
    }
    int r = check_launch("server setup");
    if (r) return r;
  }
  scf_iteration(c, it, s, S->threshold, S->max_steps, S->switch_ev, 40);
  int r = check_launch("server iteration");
  if (r) return r;
  HIP_TRY(hipMemcpyAsync(L->rb_act.p, it.active, 4 * hw, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(L->rb_steps.p, it.steps, 4 * hw, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(L->rb_conv.p, it.converged, 4 * hw, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(L->rb_stat.p, it.status, 4 * hw, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(L->rb_e.p, it.e_el, 8 * hw, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(L->rb_ecc.p, it.e_cc, 8 * hw, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipEventRecord(L->ev, s));
  L->busy = true;
  return SG_OK;
}

int server_create(sg_context* ctx, double hbm_bytes, int max_slots, int max_atoms,
                  double threshold_ev, int max_steps, int policy, double switch_ev,
                  int n_classes, const int* bounds, sg_server** out) {
  if (!ctx || !out || max_slots <= 0 || max_atoms <= 0 || !(hbm_bytes > 0) || max_steps < 0 ||
      !(threshold_ev >= 0.0) || !valid_policy(policy) ||
      n_classes < 1 || n_classes > 16 || (n_classes > 1 && !bounds) || max_slots < n_classes ||
      max_atoms < n_classes)
    return fail(SG_ERR_INVALID, "sg_server_create: bad arguments");
  for (int i = 0; i + 1 < n_classes; ++i)
    if (bounds[i] <= (i ? bounds[i - 1] : 0))
      return fail(SG_ERR_INVALID, "sg_server_create: class bounds must increase from 1");
  HIP_TRY(hipSetDevice(ctx->device));
  auto* S = new sg_server();
  S->ctx = ctx;
  S->threshold = threshold_ev;
  S->max_steps = max_steps;
  S->policy = policy;
  S->switch_ev = switch_ev;
  // per lane: pair-list scratch of 1/64 of its share of the budget, between
  // 64 Ki and 4 Mi entries, plus its slots and atoms
  const int ls = max_slots / n_classes, la = max_atoms / n_classes;
  const int64_t pc =
      std::min<int64_t>(1 << 22, std::max<int64_t>(1 << 16, (int64_t)(hbm_bytes / n_classes / 512)));
  const double fixed = (double)n_classes * (double)lane_bytes(ls, la, pc);
  // shared pools, per N^2 element of matrix space: 9 doubles (H, F, P, C, 5 for
  // the eigensolver), 2.5 pair-block elements (8 + 4 bytes), 1/8 vector element
  // (5 vectors of doubles)
  const double per = 72.0 + 2.5 * 12.0 + 5.0;
  if (hbm_bytes <= fixed + per * 1024) {
    delete S;
    return fail(SG_ERR_INVALID, "sg_server_create: HBM budget too small");
  }
  S->mat_cap = (int64_t)((hbm_bytes - fixed) / per);
  S->blk_cap = (int64_t)(2.5 * S->mat_cap);
  S->vec_cap = S->mat_cap / 8 + kMaxBasis;
  const int64_t mc = S->mat_cap, bc = S->blk_cap, vc = S->vec_cap;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_blk = take(8 * bc), o_blk32 = take(4 * bc);
  const size_t o_H = take(8 * mc), o_F = take(8 * mc), o_P = take(8 * mc), o_C = take(8 * mc),
               o_W = take(8 * 5 * mc);
  const size_t o_eig = take(8 * vc), o_dlam = take(8 * vc), o_tau = take(8 * vc),
               o_dT = take(8 * vc), o_eT = take(8 * vc);
  hipError_t e = hipMalloc(&S->pool, off);
  if (e != hipSuccess) {
    S->pool = nullptr;
    delete S;
    return fail(SG_ERR_HIP, std::string("sg_server_create: hipMalloc(") + std::to_string(off) +
                                "): " + hipGetErrorString(e));
  }
  char* base = (char*)S->pool;
  BatchDev& d = S->shared;
  d.mixed = policy == SG_POLICY_FP64 ? 0 : 1;
  d.eig32 = policy == SG_POLICY_MIXED_EIG32 ? 1 : 0;
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
  S->mat.reset(mc);
  S->blk.reset(bc);
  S->vec.reset(vc);
  for (int i = 0; i < n_classes; ++i) {
    Lane* L = new Lane();
    S->lanes.push_back(L);
    L->n_lo = i ? bounds[i - 1] : 0;
    L->n_hi = i + 1 < n_classes ? bounds[i] : kMaxBasis + 1;
    const int r = lane_create(S->shared, L, ls, la, pc,
                              i < kCtxClassStreams ? ctx->class_streams[i] : nullptr);
    if (r) {
      delete S;
      return r;
    }
  }
  S->t0 = std::chrono::steady_clock::now();
  *out = S;
  return SG_OK;
}

}  // namespace

extern "C" {

double sg_server_clock(const sg_server* S) { return S ? S->now() : 0.0; }

int sg_server_create(sg_context* ctx, double hbm_bytes, int max_slots, int max_atoms,
                     double threshold_ev, int max_steps, int policy, double switch_ev,
                     sg_server** out) {
  return server_create(ctx, hbm_bytes, max_slots, max_atoms, threshold_ev, max_steps, policy,
                       switch_ev, 1, nullptr, out);
}

int sg_server_create_classes(sg_context* ctx, double hbm_bytes, int max_slots, int max_atoms,
                             double threshold_ev, int max_steps, int policy, double switch_ev,
                             int n_classes, const int* bounds, sg_server** out) {
  return server_create(ctx, hbm_bytes, max_slots, max_atoms, threshold_ev, max_steps, policy,
                       switch_ev, n_classes, bounds, out);
}

int sg_server_set_diis(sg_server* S, int depth) {
  if (!S) return fail(SG_ERR_INVALID, "null server");
  if (depth != 0 && (depth < 2 || depth > kDiisMax))
    return fail(SG_ERR_INVALID, "DIIS depth must be 0 or lie in [2, " + std::to_string(kDiisMax) + "]");
  for (Lane* L : S->lanes) {
    bool idle = !L->busy && L->queue.empty();
    for (const Slot& sl : L->slots) idle = idle && !sl.used;
    if (!idle)
      return fail(SG_ERR_INVALID, "sg_server_set_diis: requests in flight (set the DIIS depth "
                                  "before submitting or after draining)");
  }
  HIP_TRY(hipSetDevice(S->ctx->device));
  for (Lane* L : S->lanes) HIP_TRY(hipStreamSynchronize(L->stream));
  auto clear = [](BatchDev& d) {
    d.diis_m = 0;
    d.diis_stride = 0;
    d.diis_F = d.diis_E = d.diis_B = d.diis_part = d.diis_c = nullptr;
    d.diis_k = d.diis_p = nullptr;
  };
  // back to no DIIS anywhere (also after a failed allocation below)
  auto release = [&]() {
    if (S->diis_pool) (void)hipFree(S->diis_pool);
    S->diis_pool = nullptr;
    clear(S->shared);
    for (Lane* L : S->lanes) {
      if (L->diis_dev) (void)hipFree(L->diis_dev);
      L->diis_dev = nullptr;
      clear(L->d);
    }
  };
  release();
  if (depth == 0) return SG_OK;
  const size_t hist = sizeof(double) * (size_t)depth * (size_t)S->mat_cap;
  hipError_t e = hipMalloc(&S->diis_pool, 2 * hist);
  if (e != hipSuccess) {
    S->diis_pool = nullptr;
    return fail(SG_ERR_HIP, std::string("sg_server_set_diis: hipMalloc(") +
                                std::to_string(2 * hist) + "): " + hipGetErrorString(e));
  }
  double* hF = (double*)S->diis_pool;
  double* hE = (double*)((char*)S->diis_pool + hist);
  for (Lane* L : S->lanes) {
    const size_t ns = (size_t)L->max_slots + 1;
    const size_t o_B = 0, o_part = o_B + sizeof(double) * ns * kDiisMax * kDiisMax,
                 o_c = o_part + sizeof(double) * (size_t)diis_part_doubles((int)ns),
                 o_k = o_c + sizeof(double) * ns * kDiisMax, o_p = o_k + sizeof(int) * ns,
                 end = o_p + sizeof(int) * ns;
    e = hipMalloc(&L->diis_dev, end);
    if (e != hipSuccess) {
      L->diis_dev = nullptr;
      release();
      return fail(SG_ERR_HIP, std::string("sg_server_set_diis: hipMalloc(") +
                                  std::to_string(end) + "): " + hipGetErrorString(e));
    }
    char* base = (char*)L->diis_dev;
    BatchDev& d = L->d;
    d.diis_F = hF;
    d.diis_E = hE;
    d.diis_B = (double*)(base + o_B);
    d.diis_part = (double*)(base + o_part);
    d.diis_c = (double*)(base + o_c);
    d.diis_k = (int*)(base + o_k);
    d.diis_p = (int*)(base + o_p);
    d.diis_stride = S->mat_cap;
    d.diis_m = depth;
  }
  return SG_OK;
}

int sg_server_destroy(sg_server* S) {
  if (!S) return SG_OK;
  (void)hipSetDevice(S->ctx->device);
  delete S;
  return SG_OK;
}

int sg_server_submit(sg_server* S, long long request_id, int n_atoms, const int* Z,
                     const double* xyz_ang) {
  if (!S || n_atoms <= 0 || !Z || !xyz_ang) return fail(SG_ERR_INVALID, "sg_server_submit: bad arguments");
  Request r;
  r.id = request_id;
  r.t_submit = S->now();
  r.Z.assign(Z, Z + n_atoms);
  r.xyz.assign(xyz_ang, xyz_ang + 3 * n_atoms);
  int ne = 0;
  for (int i = 0; i < n_atoms; ++i) {
    if (!supported_element(Z[i]))
      return fail(SG_ERR_UNSUPPORTED, "request " + std::to_string(request_id) + ": element Z=" +
                                          std::to_string(Z[i]) + " has no MNDO parameters here");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(xyz_ang[3 * i + k]))
        return fail(SG_ERR_INVALID, "request " + std::to_string(request_id) + ": non-finite coordinate");
    r.n_ao += n_ao_of(Z[i]);
    ne += core_charge(Z[i]);
  }
  r.n_occ = ne / 2;
  r.n_blk = pair_block_doubles(r.Z);
  Lane* L = nullptr;
  for (Lane* l : S->lanes)
    if (r.n_ao >= l->n_lo && r.n_ao < l->n_hi) L = l;
  const int64_t n2 = (int64_t)r.n_ao * r.n_ao;
  if (!L || n2 > S->mat_cap || r.n_blk > S->blk_cap || n_atoms > L->max_atoms ||
      (int64_t)n_atoms * (n_atoms - 1) + 3 * (int64_t)r.n_ao > L->pair_cap)
    return fail(SG_ERR_INVALID, "request " + std::to_string(request_id) +
                                    " does not fit this server's capacity");
  L->queue.push_back(std::move(r));
  return SG_OK;
}

int sg_server_step(sg_server* S, int* n_inflight, int* n_queued) {
  if (!S) return fail(SG_ERR_INVALID, "null server");
  HIP_TRY(hipSetDevice(S->ctx->device));
  bool moved = false;
  for (Lane* L : S->lanes) {
    if (L->busy) {
      const hipError_t e = hipEventQuery(L->ev);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) return fail(SG_ERR_HIP, std::string("server lane: ") + hipGetErrorString(e));
      int r = lane_finish(S, L);
      if (r) return r;
      moved = true;
    }
    int r = lane_launch(S, L);
    if (r) return r;
    moved = moved || L->busy;
  }
  int inflight = 0, queued = 0;
  for (Lane* L : S->lanes) {
    for (const Slot& sl : L->slots) inflight += sl.used;
    queued += (int)L->queue.size();
  }
  if (!moved && inflight > 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
  if (n_inflight) *n_inflight = inflight;
  if (n_queued) *n_queued = queued;
  return SG_OK;
}

int sg_server_poll(sg_server* S, sg_served* out, int max_out) {
  if (!S || (max_out > 0 && !out)) return fail(SG_ERR_INVALID, "sg_server_poll: bad arguments");
  const int n = std::min<int>(max_out, (int)S->done.size());
  for (int i = 0; i < n; ++i) out[i] = S->done[i];
  S->done.erase(S->done.begin(), S->done.begin() + n);
  return n;
}

}  // extern "C"
