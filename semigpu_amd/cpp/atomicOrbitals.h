// atomicOrbitals.h -- AtomicOrbital, ListAtomicOrbitals
// (reference: src/common/atomicOrbitals.h:12-77).
//
// ListAtomicOrbitals::SetOrbitals also places the molecule on the GPU (a batch
// of one molecule); every other class reaches the device through it.
#ifndef SEMIGPU_ATOMICORBITALS_H
#define SEMIGPU_ATOMICORBITALS_H

#include <cstddef>
#include <vector>
using std::vector;

#include "atom.h"
#include "semigpu_device.h"

class AtomicOrbital {
 public:
  AtomicOrbital();
  int indexAtom;
  int element;
  int angularMomentum[3];  // (0,0,0) s, (1,0,0) px, (0,1,0) py, (0,0,1) pz
  int angularMomentumInt;  // 0 s, 1 px, 2 py, 3 pz
  int indexAO;
  double coordinates[3];   // bohr
  void SetIndexAtom(int);
  void SetElement(int);
  void SetAngularMomentum(int*);
  void SetIndexAO(int);
  void SetCoordinates(const double (&coor)[3]);  // Angstrom in, bohr stored
  int GetCoreCharge() const;
  int GetAOsSize() const;
};

class ListAtomicOrbitals {
 public:
  ListAtomicOrbitals();
  ~ListAtomicOrbitals();
  ListAtomicOrbitals(const ListAtomicOrbitals&) = delete;
  ListAtomicOrbitals& operator=(const ListAtomicOrbitals&) = delete;
  AtomicOrbital* orbital;
  bool statusData;
  void SetOrbitals(const vector<Atom>&);
  size_t size() const;
  // The molecule's device batch; throws if SetOrbitals did not place it.
  sg_batch* Batch() const;
  int NumberOfAtoms() const { return nAtoms_; }

 private:
  size_t totalOrbitals_;
  int nAtoms_;
  sg_batch* batch_;
};

#endif  // SEMIGPU_ATOMICORBITALS_H
