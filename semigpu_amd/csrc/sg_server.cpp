// sg_server.cpp -- continuous-batching SCF server (C ABI in include/semigpu.h).
//
// The server owns fixed-capacity device pools carved into the same arrays the
// batch kernels use: N^2 matrix space (H, F, P, C and the eigensolver
// workspace), pair-block space (FP64 and FP32 copies), length-N vector space,
// atom descriptors and molecule slots.  Each sg_server_step()
//   1. admits queued requests in arrival order while their footprint fits the
//      free ranges (first fit) and a slot is free;
//   2. runs integrals, Hcore, core-core, guess density, guess Fock and E_0 for
//      the newly admitted molecules only (per-molecule mask);
//   3. runs one SCF iteration for every molecule in flight (eigensolver,
//      density, Fock, energy, convergence bookkeeping);
//   4. retires molecules that converged, hit the step cap or failed, frees
//      their ranges and slot, and records the result with its timestamps.
// State of molecules in flight (pair blocks, H, F, P, C) stays in HBM between
// steps.  Slot index max_slots is a permanent dead slot (never active, never
// masked): atom descriptors of freed atom ranges point to it, so kernels that
// walk the atom array skip them.  Every kernel works per molecule with a fixed summation order, so a
// molecule's result does not depend on what else is in flight.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <string>
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


}  // namespace

struct sg_server {
  sg_context* ctx = nullptr;
  int max_slots = 0, max_atoms = 0;
  int64_t mat_cap = 0, blk_cap = 0, vec_cap = 0;
  int64_t pair_cap = 0;  // scratch entries for the pair lists set up per step
  double threshold = 5e-3, switch_ev = 1e-3;
  int max_steps = 100, policy = SG_POLICY_FP64;
  void* dev = nullptr;
  BatchDev d;
  int* mask_new = nullptr;  // device, per slot (+ dead slot)
  int2* pair_scratch = nullptr;
  std::vector<int> mask;    // host copy of mask_new
  int* order_dev = nullptr;  // slots by decreasing N (device)
  std::vector<int> order;    // host copy
  bool order_dirty = true;
  std::set<int> free_slots;
  std::vector<MolDesc> mols;
  std::vector<AtomDesc> atoms;
  std::vector<Slot> slots;
  std::deque<Request> queue;
  RangeAlloc mat, blk, vec, atom;
  std::vector<sg_served> done;
  std::chrono::steady_clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

extern "C" {

double sg_server_clock(const sg_server* S) { return S ? S->now() : 0.0; }

int sg_server_create(sg_context* ctx, double hbm_bytes, int max_slots, int max_atoms,
                     double threshold_ev, int max_steps, int policy, double switch_ev,
                     sg_server** out) {
  if (!ctx || !out || max_slots <= 0 || max_atoms <= 0 || !(hbm_bytes > 0) || max_steps < 0 ||
      !(threshold_ev >= 0.0) || (policy != SG_POLICY_FP64 && policy != SG_POLICY_MIXED))
    return fail(SG_ERR_INVALID, "sg_server_create: bad arguments");
  auto* S = new sg_server();
  S->ctx = ctx;
  S->max_slots = max_slots;
  S->max_atoms = max_atoms;
  S->threshold = threshold_ev;
  S->max_steps = max_steps;
  S->policy = policy;
  S->switch_ev = switch_ev;
  // Budget split per N^2 element of matrix space: 9 doubles (H, F, P, C, 5 for
  // the eigensolver), 2.5 pair-block elements (8 + 4 bytes), 1/8 vector element
  // (4 vectors of doubles).
  // pair-list scratch: 1/64 of the budget, between 64 Ki and 4 Mi entries
  S->pair_cap = std::min<int64_t>(1 << 22, std::max<int64_t>(1 << 16, (int64_t)(hbm_bytes / 512)));
  const double fixed = (double)(max_slots + 1) * (sizeof(MolDesc) + 16 * 8) +
                       (double)max_atoms * sizeof(AtomDesc) + (double)S->pair_cap * sizeof(int2);
  const double per = 72.0 + 2.5 * 12.0 + 4.0;
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
  const size_t ns = max_slots + 1;  // + dead slot
  const size_t o_mols = take(sizeof(MolDesc) * ns), o_atoms = take(sizeof(AtomDesc) * max_atoms);
  const size_t o_pairs = take(sizeof(int2) * S->pair_cap);
  const size_t o_blk = take(8 * bc), o_blk32 = take(4 * bc);
  const size_t o_H = take(8 * mc), o_F = take(8 * mc), o_P = take(8 * mc), o_C = take(8 * mc),
               o_W = take(8 * 5 * mc);
  const size_t o_eig = take(8 * vc), o_tau = take(8 * vc), o_dT = take(8 * vc), o_eT = take(8 * vc);
  const size_t o_tn = take(8 * ns), o_ecc = take(8 * ns), o_eel = take(8 * ns), o_enew = take(8 * ns);
  const size_t o_steps = take(4 * ns), o_act = take(4 * ns), o_conv = take(4 * ns),
               o_stat = take(4 * ns), o_nact = take(256), o_prec = take(4 * ns),
               o_pprec = take(4 * ns), o_mask = take(4 * ns), o_order = take(4 * ns);
  (void)hipSetDevice(ctx->device);
  hipError_t e = hipMalloc(&S->dev, off);
  if (e != hipSuccess) {
    delete S;
    return fail(SG_ERR_HIP, std::string("sg_server_create: hipMalloc(") + std::to_string(off) +
                                "): " + hipGetErrorString(e));
  }
  char* base = (char*)S->dev;
  BatchDev& d = S->d;
  d.mols = (MolDesc*)(base + o_mols);
  d.atoms = (AtomDesc*)(base + o_atoms);
  S->pair_scratch = (int2*)(base + o_pairs);
  d.blocks = (double*)(base + o_blk);
  d.blocks32 = (float*)(base + o_blk32);
  d.H = (double*)(base + o_H);
  d.F = (double*)(base + o_F);
  d.P = (double*)(base + o_P);
  d.C = (double*)(base + o_C);
  d.W = (double*)(base + o_W);
  d.eig = (double*)(base + o_eig);
  d.tau = (double*)(base + o_tau);
  d.dT = (double*)(base + o_dT);
  d.eT = (double*)(base + o_eT);
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
  S->mask_new = (int*)(base + o_mask);
  S->order_dev = (int*)(base + o_order);
  hipStream_t s = ctx->stream;
  for (int* p : {d.steps, d.active, d.converged, d.status, d.prec, d.prev_prec, S->mask_new})
    if (e == hipSuccess) e = hipMemsetAsync(p, 0, 4 * ns, s);
  S->mols.assign(ns, MolDesc{});
  AtomDesc dead{};
  dead.mol = max_slots;
  S->atoms.assign(max_atoms, dead);
  if (e == hipSuccess) e = hipMemcpyAsync(d.mols, S->mols.data(), sizeof(MolDesc) * ns, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d.atoms, S->atoms.data(), sizeof(AtomDesc) * max_atoms, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(S->dev);
    delete S;
    return fail(SG_ERR_HIP, std::string("sg_server_create: ") + hipGetErrorString(e));
  }
  S->slots.assign(max_slots, Slot{});
  S->mask.assign(ns, 0);
  for (int i = 0; i < max_slots; ++i) S->free_slots.insert(i);
  S->mat.reset(mc);
  S->blk.reset(bc);
  S->vec.reset(vc);
  S->atom.reset(max_atoms);
  S->t0 = std::chrono::steady_clock::now();
  *out = S;
  return SG_OK;
}

int sg_server_destroy(sg_server* S) {
  if (!S) return SG_OK;
  (void)hipSetDevice(S->ctx->device);
  (void)hipStreamSynchronize(S->ctx->stream);
  if (S->dev) (void)hipFree(S->dev);
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
  const int64_t n2 = (int64_t)r.n_ao * r.n_ao;
  if (r.n_ao > kMaxBasis || n2 > S->mat_cap || r.n_blk > S->blk_cap || n_atoms > S->max_atoms ||
      (int64_t)n_atoms * (n_atoms - 1) + 3 * (int64_t)r.n_ao > S->pair_cap)
    return fail(SG_ERR_INVALID, "request " + std::to_string(request_id) +
                                    " does not fit this server's capacity");
  S->queue.push_back(std::move(r));
  return SG_OK;
}

int sg_server_step(sg_server* S, int* n_inflight, int* n_queued) {
  if (!S) return fail(SG_ERR_INVALID, "null server");
  sg_context* c = S->ctx;
  hipStream_t s = c->stream;
  HIP_TRY(hipSetDevice(c->device));
  // 1. admission in arrival order, first fit
  std::vector<int> fresh;
  std::vector<int2> pxx, phx, phh, pall, diag;
  int64_t pairs_budget = S->pair_cap;
  while (!S->queue.empty()) {
    Request& r = S->queue.front();
    const int na = (int)r.Z.size();
    // scratch entries: each atom pair twice (type list + full list), <= 10
    // Hcore tasks per heavy atom (4 AOs) and 1 per hydrogen, i.e. < 3 N
    const int64_t need = (int64_t)na * (na - 1) + 3 * (int64_t)r.n_ao;
    if (need > pairs_budget || S->free_slots.empty()) break;
    const int slot = *S->free_slots.begin();
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
    if (!S->atom.alloc(na, &a0)) {
      S->mat.release(m0, n2);
      S->blk.release(b0, r.n_blk);
      S->vec.release(v0, r.n_ao);
      break;
    }
    S->free_slots.erase(S->free_slots.begin());
    Slot& sl = S->slots[slot];
    sl.used = true;
    sl.t_admit = S->now();
    sl.mat0 = m0;
    sl.blk0 = b0;
    sl.vec0 = v0;
    sl.atom0 = a0;
    MolDesc& M = S->mols[slot];
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
      AtomDesc& A = S->atoms[a0 + a];
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
    sl.req = std::move(r);
    S->queue.pop_front();
    fresh.push_back(slot);
    pairs_budget -= need;
  }
  // high-water marks of the slot and atom ranges in use
  int hw = 0, hw_atoms = 0, max_n = 1;
  for (int i = 0; i < S->max_slots; ++i)
    if (S->slots[i].used) {
      hw = i + 1;
      hw_atoms = std::max(hw_atoms, S->mols[i].atom0 + S->mols[i].n_atoms);
      max_n = std::max(max_n, S->mols[i].n_ao);
    }
  BatchDev it = S->d;
  it.n_mol = hw;
  it.n_atoms = hw_atoms;
  it.max_n = max_n;
  if (!fresh.empty()) S->order_dirty = true;
  if (S->order_dirty && hw > 0) {
    // slots in flight by decreasing N, then the free slots below the high-water mark
    S->order.resize(hw);
    for (int i = 0; i < hw; ++i) S->order[i] = i;
    std::stable_sort(S->order.begin(), S->order.end(), [&](int x, int y) {
      const int nx = S->slots[x].used ? S->mols[x].n_ao : -1;
      const int ny = S->slots[y].used ? S->mols[y].n_ao : -1;
      return nx > ny;
    });
    HIP_TRY(hipMemcpyAsync(S->order_dev, S->order.data(), sizeof(int) * hw, hipMemcpyHostToDevice, s));
    S->order_dirty = false;
  }
  it.order = S->order_dev;
  if (!fresh.empty()) {
    for (int sl : fresh) S->mask[sl] = 1;
    HIP_TRY(hipMemcpyAsync(S->d.mols, S->mols.data(), sizeof(MolDesc) * hw, hipMemcpyHostToDevice, s));
    for (int sl : fresh) {
      const MolDesc& M = S->mols[sl];
      HIP_TRY(hipMemcpyAsync(S->d.atoms + M.atom0, S->atoms.data() + M.atom0,
                             sizeof(AtomDesc) * M.n_atoms, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipMemcpyAsync(S->mask_new, S->mask.data(), sizeof(int) * S->mask.size(),
                           hipMemcpyHostToDevice, s));
    // pair lists and Hcore tasks of the new molecules, packed into the scratch
    BatchDev nb = it;
    int2* p = S->pair_scratch;
    auto put = [&](const std::vector<int2>& v, int2*& dst, int& cnt) -> hipError_t {
      dst = p;
      cnt = (int)v.size();
      p += v.size();
      return v.empty() ? hipSuccess
                       : hipMemcpyAsync(dst, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice, s);
    };
    HIP_TRY(put(pxx, nb.pairs_xx, nb.n_xx));
    HIP_TRY(put(phx, nb.pairs_hx, nb.n_hx));
    HIP_TRY(put(phh, nb.pairs_hh, nb.n_hh));
    HIP_TRY(put(pall, nb.pairs_all, nb.n_pairs_all));
    HIP_TRY(put(diag, nb.diag_tasks, nb.n_diag_tasks));
    {
      StageTimer t(c, ST_TWO_CENTER);
      launch_pair_integrals(nb, s, S->mask_new);
    }
    {
      StageTimer t(c, ST_HCORE);
      launch_clear_hf(nb, s, S->mask_new);
      launch_hcore(nb, s);
      launch_core_core(nb, s, S->mask_new);
    }
    {
      StageTimer t(c, ST_GUESS);
      launch_guess_density(nb, s, S->mask_new);
    }
    {
      StageTimer t(c, ST_FOCK);
      launch_fock(nb, s, true, S->mask_new);
    }
    {
      StageTimer t(c, ST_ENERGY);
      launch_energy(nb, s, nb.e_new, true, S->mask_new);
    }
    {
      StageTimer t(c, ST_CONTROL);
      launch_set_prec(nb, s, S->policy == SG_POLICY_MIXED ? 1 : 0, S->mask_new);
      launch_scf_init(nb, s, S->max_steps, S->policy, S->mask_new);
    }
    int r = check_launch("server setup");
    if (r) return r;
  }
  // 2. one SCF iteration for every molecule in flight
  if (hw > 0) {
    scf_iteration(c, it, s, S->threshold, S->max_steps, S->switch_ev, 40);
    int r = check_launch("server iteration");
    if (r) return r;
    std::vector<int> act(hw), st(hw), cv(hw), ss(hw);
    std::vector<double> e(hw), ecc(hw);
    HIP_TRY(hipMemcpyAsync(act.data(), it.active, 4 * hw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(st.data(), it.steps, 4 * hw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(cv.data(), it.converged, 4 * hw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ss.data(), it.status, 4 * hw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(e.data(), it.e_el, 8 * hw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ecc.data(), it.e_cc, 8 * hw, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // host buffers read by this step's uploads are free again
    for (int sl : fresh) S->mask[sl] = 0;
    const double tdone = S->now();
    // 3. retire finished molecules
    bool retired = false;
    for (int i = 0; i < hw; ++i) {
      Slot& sl = S->slots[i];
      if (!sl.used || act[i]) continue;
      sg_served o;
      memset(&o, 0, sizeof(o));
      o.request_id = sl.req.id;
      o.e_el = e[i];
      o.e_cc = ecc[i];
      o.e_tot = e[i] + ecc[i];
      o.steps = st[i];
      o.converged = cv[i];
      o.status = ss[i];
      o.n_ao = sl.req.n_ao;
      o.t_submit = sl.req.t_submit;
      o.t_admit = sl.t_admit;
      o.t_done = tdone;
      S->done.push_back(o);
      const MolDesc& M = S->mols[i];
      S->mat.release(sl.mat0, (int64_t)M.n_ao * M.n_ao);
      S->blk.release(sl.blk0, sl.req.n_blk);
      S->vec.release(sl.vec0, M.n_ao);
      S->atom.release(sl.atom0, M.n_atoms);
      AtomDesc dead{};
      dead.mol = S->max_slots;
      std::fill(S->atoms.begin() + M.atom0, S->atoms.begin() + M.atom0 + M.n_atoms, dead);
      HIP_TRY(hipMemcpyAsync(S->d.atoms + M.atom0, S->atoms.data() + M.atom0,
                             sizeof(AtomDesc) * M.n_atoms, hipMemcpyHostToDevice, s));
      S->free_slots.insert(i);
      sl = Slot{};
      retired = true;
      S->order_dirty = true;
    }
    // the host descriptors must outlive the copies above
    if (retired) HIP_TRY(hipStreamSynchronize(s));
  }
  int inflight = 0;
  for (const Slot& sl : S->slots) inflight += sl.used;
  if (n_inflight) *n_inflight = inflight;
  if (n_queued) *n_queued = (int)S->queue.size();
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
