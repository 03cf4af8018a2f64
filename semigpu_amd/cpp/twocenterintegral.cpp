// TwoCenterIntegral (twocenterintegral.cpp:18-386 of the reference).
#include "twocenterintegral.h"

#include <cstring>
#include <map>
#include <mutex>

namespace {

// MOPAC index of the AO pair (l1, l2), l = 0 s, 1 px, 2 py, 3 pz
// (GetValueFromArray, twocenterintegral.cpp:307-318).
int mopac_pair(int a, int b) {
  if (a > b) std::swap(a, b);
  return b * (b + 1) / 2 + a;
}

int block_rows(int zi, int zj) { return (zi > 1 && zj > 1) ? 10 : 1; }
int block_cols(int zi, int zj) { return (zi == 1 && zj == 1) ? 1 : 10; }

// Which device molecule each host array was filled from (for To_device).
struct Owner {
  sg_batch* batch;
  vector<int> Z;
};
std::map<double****, Owner>& owners() {
  static std::map<double****, Owner> m;
  return m;
}
std::mutex& owners_mu() {
  static std::mutex m;
  return m;
}

// Walks the jagged array in the flat reference order [i][j<=i][k][l].
template <class F>
void for_each_block(const vector<int>& Z, double**** a, F f) {
  const int n = (int)Z.size();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const int r = block_rows(Z[i], Z[j]), c = block_cols(Z[i], Z[j]);
      for (int k = 0; k < r; ++k) f(a[i][j][k], c);
    }
}

vector<int> atom_numbers(const vector<Atom>& m) {
  vector<int> Z;
  for (const Atom& a : m) Z.push_back(a.atomNumber);
  return Z;
}

}  // namespace

TwoCenterIntegral::TwoCenterIntegral(MNDOparameter&) : keyZ_{0, 0}, keyXyz_{}, keyValid_(false) {}

TwoCenterIntegral::~TwoCenterIntegral() {}

bool TwoCenterIntegral::Alloc4AllTwoCenterIntegral(const vector<Atom>& molecule,
                                                   double****& a) {
  const int n = (int)molecule.size();
  if (n == 0) return false;
  a = new double***[n];
  for (int i = 0; i < n; ++i) {
    a[i] = new double**[i + 1];
    for (int j = 0; j <= i; ++j) {
      const int zi = molecule[i].atomNumber, zj = molecule[j].atomNumber;
      const int r = block_rows(zi, zj), c = block_cols(zi, zj);
      a[i][j] = new double*[r];
      double* data = new double[(size_t)r * c]();
      for (int k = 0; k < r; ++k) a[i][j][k] = data + (size_t)k * c;
    }
  }
  return true;
}

bool TwoCenterIntegral::Dealloc4AllTwoCenterIntegral(const vector<Atom>& molecule,
                                                     double****& a) {
  if (!a) return false;
  {
    std::lock_guard<std::mutex> g(owners_mu());
    owners().erase(a);
  }
  const int n = (int)molecule.size();
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j <= i; ++j) {
      delete[] a[i][j][0];
      delete[] a[i][j];
    }
    delete[] a[i];
  }
  delete[] a;
  a = nullptr;
  return true;
}

void TwoCenterIntegral::ComputeAllTwoCenterIntegral(const ListAtomicOrbitals& infoAOs,
                                                    double****& a) {
  sg_batch* b = infoAOs.Batch();
  semigpu::check(sg_two_center(b), "sg_two_center");
  const long n = sg_pair_block_count(b, 0);
  vector<double> flat(n);
  semigpu::check(sg_get_pair_blocks(b, 0, flat.data()), "sg_get_pair_blocks");
  vector<int> Z(infoAOs.NumberOfAtoms());
  for (size_t k = 0; k < infoAOs.size(); ++k) Z[infoAOs.orbital[k].indexAtom] = infoAOs.orbital[k].element;
  size_t pos = 0;
  for_each_block(Z, a, [&](double* row, int c) {
    std::memcpy(row, flat.data() + pos, sizeof(double) * c);
    pos += c;
  });
  std::lock_guard<std::mutex> g(owners_mu());
  owners()[a] = Owner{b, Z};
}

void TwoCenterIntegral::To_device(const vector<Atom>& molecule, double****& a) {
  Owner o;
  {
    std::lock_guard<std::mutex> g(owners_mu());
    auto it = owners().find(a);
    if (it == owners().end()) return;
    o = it->second;
  }
  if (o.Z != atom_numbers(molecule)) throw std::runtime_error("To_device: molecule mismatch");
  vector<double> flat;
  for_each_block(o.Z, a, [&](double* row, int c) { flat.insert(flat.end(), row, row + c); });
  semigpu::check(sg_set_pair_blocks(o.batch, 0, flat.data()), "sg_set_pair_blocks");
}

double TwoCenterIntegral::GetValueFromArray(const AtomicOrbital& A, const AtomicOrbital& B,
                                            const AtomicOrbital& C, const AtomicOrbital& D,
                                            double**** a) {
  if (A.indexAtom != B.indexAtom || C.indexAtom != D.indexAtom) return 0.0;
  // the block is stored once, under the higher atom index
  const bool swap = A.indexAtom < C.indexAtom;
  const AtomicOrbital& hi1 = swap ? C : A;
  const AtomicOrbital& hi2 = swap ? D : B;
  const AtomicOrbital& lo1 = swap ? A : C;
  const AtomicOrbital& lo2 = swap ? B : D;
  double** blk = a[hi1.indexAtom][lo1.indexAtom];
  const bool hiX = hi1.element > 1, loX = lo1.element > 1;
  if (hiX && loX)
    return blk[mopac_pair(hi1.angularMomentumInt, hi2.angularMomentumInt)]
              [mopac_pair(lo1.angularMomentumInt, lo2.angularMomentumInt)];
  if (hiX) return blk[0][mopac_pair(hi1.angularMomentumInt, hi2.angularMomentumInt)];
  if (loX) return blk[0][mopac_pair(lo1.angularMomentumInt, lo2.angularMomentumInt)];
  return blk[0][0];
}

double TwoCenterIntegral::ComputeTwoCenterIntegral(const AtomicOrbital& A,
                                                   const AtomicOrbital& B,
                                                   const AtomicOrbital& C,
                                                   const AtomicOrbital& D) {
  if (A.indexAtom != B.indexAtom || C.indexAtom != D.indexAtom) return 0.0;
  // A one- or two-atom molecule with the atoms in their original order, so
  // that its block is the one the full molecule stores.
  const bool one = A.indexAtom == C.indexAtom;
  const AtomicOrbital& lo = A.indexAtom <= C.indexAtom ? A : C;
  const AtomicOrbital& hi = A.indexAtom <= C.indexAtom ? C : A;
  const int Z[2] = {lo.element, hi.element};
  const double xyz[6] = {lo.coordinates[0], lo.coordinates[1], lo.coordinates[2],
                         hi.coordinates[0], hi.coordinates[1], hi.coordinates[2]};
  const int na = one ? 1 : 2;
  const bool same = keyValid_ && keyZ_[0] == Z[0] && keyZ_[1] == (one ? 0 : Z[1]) &&
                    std::memcmp(keyXyz_, xyz, sizeof(double) * 3 * na) == 0;
  if (!same) {
    int off[2] = {0, na};
    sg_batch* b = nullptr;
    semigpu::check(sg_batch_create_bohr(semigpu::context(), 1, off, Z, xyz, &b),
                   "sg_batch_create_bohr");
    int rc = sg_two_center(b);
    if (rc == SG_OK) {
      pairBlocks_.assign(sg_pair_block_count(b, 0), 0.0);
      rc = sg_get_pair_blocks(b, 0, pairBlocks_.data());
    }
    sg_batch_destroy(b);
    semigpu::check(rc, "two-center block");
    keyZ_[0] = Z[0];
    keyZ_[1] = one ? 0 : Z[1];
    std::memcpy(keyXyz_, xyz, sizeof(double) * 3 * na);
    keyValid_ = true;
  }
  // flat blocks of the small molecule: (0,0) [, (1,0), (1,1)]
  const int s00 = block_rows(Z[0], Z[0]) * block_cols(Z[0], Z[0]);
  const double* blk = pairBlocks_.data();
  if (one) {
    const int p = mopac_pair(A.angularMomentumInt, B.angularMomentumInt);
    const int q = mopac_pair(C.angularMomentumInt, D.angularMomentumInt);
    return Z[0] > 1 ? blk[10 * p + q] : blk[0];
  }
  blk += s00;  // block (1, 0): row = pair on the later atom
  const AtomicOrbital& h1 = A.indexAtom > C.indexAtom ? A : C;
  const AtomicOrbital& h2 = A.indexAtom > C.indexAtom ? B : D;
  const AtomicOrbital& l1 = A.indexAtom > C.indexAtom ? C : A;
  const AtomicOrbital& l2 = A.indexAtom > C.indexAtom ? D : B;
  const bool hX = Z[1] > 1, lX = Z[0] > 1;
  if (hX && lX)
    return blk[10 * mopac_pair(h1.angularMomentumInt, h2.angularMomentumInt) +
               mopac_pair(l1.angularMomentumInt, l2.angularMomentumInt)];
  if (hX) return blk[mopac_pair(h1.angularMomentumInt, h2.angularMomentumInt)];
  if (lX) return blk[mopac_pair(l1.angularMomentumInt, l2.angularMomentumInt)];
  return blk[0];
}
