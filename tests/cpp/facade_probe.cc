// Test driver for the C++ classes with the reference's names: runs the
// reference's call sequence (semiGPU.cc:20-116) on one XYZ file and writes the
// numbers the GPU tests compare with the fixtures.
//
//   facade_probe file.xyz out.txt [threshold]
//
// out.txt: "key value..." lines -- n_ao, hcore (packed), e_el, e_cc, steps,
// converged, and the brute-force check of ComputeTwoCenterIntegral against
// GetValueFromArray over every AO quadruple (test_twoCenterIntegral.cc:289-313
// of the reference does the same on its molecules).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "atomicOrbitals.h"
#include "corecorerepulsion.h"
#include "hcore.h"
#include "readxyzfile.h"
#include "scfcalculation.h"
#include "twocenterintegral.h"

int main(int argc, char* argv[]) {
  if (argc < 3) return 2;
  try {
    ReadXYZFile reader;
    vector<Atom> mol;
    if (!reader.GetValuesFromFile(argv[1], mol)) return 3;
    ListAtomicOrbitals aos;
    aos.SetOrbitals(mol);
    if (!aos.statusData) return 4;
    double**** ints = nullptr;
    TwoCenterIntegral::Alloc4AllTwoCenterIntegral(mol, ints);
    MNDOparameter par;
    TwoCenterIntegral tci(par);
    tci.ComputeAllTwoCenterIntegral(aos, ints);
    TwoCenterIntegral::To_device(mol, ints);
    const size_t n = aos.size();
    // brute force: every quadruple with mu, nu on one atom and lambda, sigma on one atom
    long checked = 0, mismatched = 0, zero_bad = 0;
    double worst = 0.0;
    for (size_t a = 0; a < n; ++a)
      for (size_t b = 0; b < n; ++b) {
        const AtomicOrbital &A = aos.orbital[a], &B = aos.orbital[b];
        for (size_t c = 0; c < n; ++c)
          for (size_t d = 0; d < n; ++d) {
            const AtomicOrbital &C = aos.orbital[c], &D = aos.orbital[d];
            const double g = TwoCenterIntegral::GetValueFromArray(A, B, C, D, ints);
            if (A.indexAtom != B.indexAtom || C.indexAtom != D.indexAtom) {
              zero_bad += g != 0.0;
              continue;
            }
            const double v = tci.ComputeTwoCenterIntegral(A, B, C, D);
            ++checked;
            if (v != g) {
              ++mismatched;
              worst = std::fmax(worst, std::fabs(v - g));
            }
          }
      }
    Hcore h(par, aos, ints);
    h.Alloc4Matrix("hcoreMatrix");
    h.ComputeMatrix();
    SCFCalculation scf(aos, ints, h);
    if (argc > 3) scf.thresholdEnergy = std::atof(argv[3]);
    scf.AllocSCFData();
    scf.ComputeSCF();
    const double ecc = CoreCoreRepulsion::ComputeRepulsion(par, aos, ints);
    std::ofstream out(argv[2]);
    out << std::setprecision(17);
    out << "n_ao " << n << "\n";
    out << "quadruples " << checked << " " << mismatched << " " << worst << " " << zero_bad << "\n";
    out << "hcore";
    for (size_t i = 0; i < h.array1DSize_; ++i) out << " " << h.matrixHold_[i];
    out << "\n";
    out << "element";  // ComputeElementMatrix(i, j) for the last row
    for (size_t j = 0; j < n; ++j) out << " " << h.ComputeElementMatrix(n - 1, j);
    out << "\n";
    out << "e_el " << scf.finalEnergy << "\n";
    out << "e_cc " << ecc << "\n";
    out << "converged " << (scf.energyConverge ? 1 : 0) << "\n";
    out << "eigenvalues";
    for (size_t i = 0; i < n; ++i) out << " " << scf.eigenVal[i];
    out << "\n";
    out << "density";
    for (size_t i = 0; i < scf.Pmatrix->array1DSize_; ++i) out << " " << scf.Pmatrix->matrixHold_[i];
    out << "\n";
    scf.DeallocSCFData();
    h.Dealloc4Matrix();
    TwoCenterIntegral::Dealloc4AllTwoCenterIntegral(mol, ints);
  } catch (const std::exception& e) {
    std::cerr << "facade_probe: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
