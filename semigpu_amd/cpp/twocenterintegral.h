// twocenterintegral.h -- TwoCenterIntegral (reference: src/common/twocenterintegral.h:14-33).
//
// double**** all2CenterIntegral keeps the reference's jagged host layout
// [i][j<=i][k][l]: 1x1 for H-H, 1x10 for H-X / X-H (k = 0, l = the pair on
// the heavy atom), 10x10 for X-X, pairs in MOPAC order SS,SX,XX,SY,XY,YY,SZ,XZ,YZ,ZZ.
// The integrals are computed on the GPU and copied into it; the device copy
// is the one the Hcore, Fock and core-core stages read.
#ifndef SEMIGPU_TWOCENTERINTEGRAL_H
#define SEMIGPU_TWOCENTERINTEGRAL_H

#include <vector>
using std::vector;

#include "MNDO_parameters.h"
#include "atomicOrbitals.h"

class TwoCenterIntegral {
 public:
  TwoCenterIntegral(MNDOparameter&);
  ~TwoCenterIntegral();
  // One (mu nu | lambda sigma) in eV; 0 unless mu, nu share an atom and
  // lambda, sigma share an atom.
  double ComputeTwoCenterIntegral(const AtomicOrbital& orbitalA, const AtomicOrbital& orbitalB,
                                  const AtomicOrbital& orbitalC, const AtomicOrbital& orbitalD);
  static bool Alloc4AllTwoCenterIntegral(const vector<Atom>& molecule,
                                         double****& all2CenterIntegral);
  static bool Dealloc4AllTwoCenterIntegral(const vector<Atom>& molecule,
                                           double****& all2CenterIntegral);
  // Computes every block on the device and copies them into all2CenterIntegral.
  void ComputeAllTwoCenterIntegral(const ListAtomicOrbitals& infoAOs,
                                   double****& all2CenterIntegral);
  static double GetValueFromArray(const AtomicOrbital&, const AtomicOrbital&,
                                  const AtomicOrbital&, const AtomicOrbital&,
                                  double**** all2CenterIntegral);
  // Copies the host blocks to the device of the molecule they were computed
  // for (after host-side edits); a no-op for arrays never computed here.
  static void To_device(const vector<Atom>& molecule, double****& all2CenterIntegral);

 private:
  // blocks of the last atom pair asked of ComputeTwoCenterIntegral
  int keyZ_[2];
  double keyXyz_[6];
  bool keyValid_;
  vector<double> pairBlocks_;
};

#endif  // SEMIGPU_TWOCENTERINTEGRAL_H
