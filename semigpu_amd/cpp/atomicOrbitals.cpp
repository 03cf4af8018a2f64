// AtomicOrbital / ListAtomicOrbitals (atomicOrbitals.cpp:10-119 of the reference).
#include "atomicOrbitals.h"

#include <iostream>

AtomicOrbital::AtomicOrbital()
    : indexAtom(0), element(0), angularMomentum{0, 0, 0}, angularMomentumInt(0), indexAO(0),
      coordinates{0.0, 0.0, 0.0} {}

void AtomicOrbital::SetIndexAtom(int v) { indexAtom = v; }
void AtomicOrbital::SetElement(int v) { element = v; }
void AtomicOrbital::SetIndexAO(int v) { indexAO = v; }

void AtomicOrbital::SetAngularMomentum(int* l) {
  angularMomentumInt = 0;
  for (int i = 0; i < 3; ++i) {
    angularMomentum[i] = l[i];
    if (l[i] == 1) angularMomentumInt = i + 1;
  }
}

void AtomicOrbital::SetCoordinates(const double (&c)[3]) {
  for (int i = 0; i < 3; ++i) coordinates[i] = c[i] * 1.8897259886;  // mymath.h Angstrom -> bohr
}

int AtomicOrbital::GetCoreCharge() const {
  if (element < 2) return element;
  if (element < 11) return element - 2;
  if (element < 19) return element - 10;
  return -1;
}

int AtomicOrbital::GetAOsSize() const {
  if (element < 2) return 1;
  if (element < 11) return 4;
  return -1;
}

ListAtomicOrbitals::ListAtomicOrbitals()
    : orbital(nullptr), statusData(true), totalOrbitals_(0), nAtoms_(0), batch_(nullptr) {}

ListAtomicOrbitals::~ListAtomicOrbitals() {
  if (batch_) sg_batch_destroy(batch_);
  delete[] orbital;
}

void ListAtomicOrbitals::SetOrbitals(const vector<Atom>& molecule) {
  if (batch_) sg_batch_destroy(batch_);
  batch_ = nullptr;
  delete[] orbital;
  nAtoms_ = (int)molecule.size();
  size_t n = 0;
  for (const Atom& a : molecule) {
    if (a.atomNumber > 18 || a.atomNumber < 1) statusData = false;
    n += a.atomNumber <= 2 ? 1 : 4;
  }
  totalOrbitals_ = n;
  orbital = new AtomicOrbital[n];
  int k = 0;
  for (int i = 0; i < nAtoms_; ++i) {
    const Atom& a = molecule[i];
    const int count = a.atomNumber <= 2 ? 1 : 4;
    for (int t = 0; t < count; ++t, ++k) {
      int l[3] = {t == 1, t == 2, t == 3};
      orbital[k].SetIndexAtom(i);
      orbital[k].SetElement(a.atomNumber);
      orbital[k].SetAngularMomentum(l);
      orbital[k].SetIndexAO(k);
      orbital[k].SetCoordinates(a.atomCoordinates);
    }
  }
  // the molecule on the device: a batch of one
  std::vector<int> Z(nAtoms_), off = {0, nAtoms_};
  std::vector<double> xyz(3 * nAtoms_);
  for (int i = 0; i < nAtoms_; ++i) {
    Z[i] = molecule[i].atomNumber;
    for (int c = 0; c < 3; ++c) xyz[3 * i + c] = molecule[i].atomCoordinates[c];
  }
  if (nAtoms_ == 0 || sg_batch_create(semigpu::context(), 1, off.data(), Z.data(), xyz.data(),
                                      &batch_) != SG_OK) {
    if (nAtoms_) std::cout << "semigpu: " << sg_last_error() << std::endl;
    batch_ = nullptr;
    statusData = false;
  }
}

size_t ListAtomicOrbitals::size() const { return totalOrbitals_; }

sg_batch* ListAtomicOrbitals::Batch() const {
  if (!batch_)
    throw std::runtime_error("ListAtomicOrbitals: no device molecule (SetOrbitals failed or not called)");
  return batch_;
}
