// Test-only harness: drives the reference CPU path (compiled from the sources
// under the reference tree, never copied) and dumps every intermediate of the
// MNDO SCF so the C restatement in mndo_oracle.c can be checked against it.
//
// Usage: ref_dump <geometry.xyz> <out.bin> <threshold_eV> [max_steps]
//
// The SCF loop below follows SCFCalculation::ComputeSCF
// (src/common/scfcalculation.cpp:66-146) with the threshold taken from the
// command line instead of the hard-coded 5e-3 eV; it calls the reference's
// own InitialGuess / GetEigenValVecOfFmatrix / ComputeMatrix / ComputeEnergy,
// so only the loop control is ours.  The shipped ComputeSCF is also run once
// (threshold 5e-3) so a test can check that both drivers agree.
//
// Output: a sequence of records [int32 name_len][name][int64 count][count doubles].
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <iostream>
#include <fstream>
#include <sstream>
#include <iomanip>
#include <chrono>

#include "readxyzfile.h"
#include "atom.h"
#include "atomicOrbitals.h"
#include "MNDO_parameters.h"
#include "twocenterintegral.h"
#include "hcore.h"
#include "densitymatrix.h"
#include "fockmatrix.h"
#include "electronicenergy.h"
#include "corecorerepulsion.h"
#include "mymemory.h"
#include "mymath.h"
// Access to the SCF driver's private helpers for the threshold-parameterised loop.
#define private public
#include "scfcalculation.h"
#undef private

static FILE* g_out = nullptr;

static void put(const char* name, const double* v, long long n) {
  int len = (int)strlen(name);
  fwrite(&len, sizeof(int), 1, g_out);
  fwrite(name, 1, len, g_out);
  fwrite(&n, sizeof(long long), 1, g_out);
  if (n) fwrite(v, sizeof(double), n, g_out);
}
static void put1(const char* name, double v) { put(name, &v, 1); }

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s geometry.xyz out.bin threshold [max_steps]\n", argv[0]);
    return 2;
  }
  std::string xyz = argv[1];
  double threshold = atof(argv[3]);
  unsigned int maxSteps = argc > 4 ? (unsigned)atoi(argv[4]) : 100u;
  g_out = fopen(argv[2], "wb");
  if (!g_out) { perror("open out"); return 2; }

  // Silence the reference's stdout chatter.
  std::stringstream sink;
  std::streambuf* old = std::cout.rdbuf(sink.rdbuf());

  ReadXYZFile reader;
  vector<Atom> molecule;
  if (!reader.GetValuesFromFile(xyz, molecule)) {
    std::cout.rdbuf(old);
    fprintf(stderr, "cannot read %s\n", xyz.c_str());
    return 1;
  }
  int nat = (int)molecule.size();
  {
    std::vector<double> z(nat), c(3 * nat);
    for (int i = 0; i < nat; ++i) {
      z[i] = molecule[i].atomNumber;
      for (int k = 0; k < 3; ++k) c[3 * i + k] = molecule[i].atomCoordinates[k];
    }
    put("Z", z.data(), nat);
    put("xyz_ang", c.data(), 3 * nat);
  }
  ListAtomicOrbitals infoAOs;
  infoAOs.SetOrbitals(molecule);
  size_t N = infoAOs.size();
  put1("n_ao", (double)N);

  double**** ints = NULL;
  TwoCenterIntegral::Alloc4AllTwoCenterIntegral(molecule, ints);
  MNDOparameter parameters;
  TwoCenterIntegral tci(parameters);
  auto t0 = std::chrono::high_resolution_clock::now();
  tci.ComputeAllTwoCenterIntegral(infoAOs, ints);
  auto t1 = std::chrono::high_resolution_clock::now();

  // Pair blocks in reference order [i][j<=i][k][l], sizes 1 / 10 / 100.
  {
    std::vector<double> blocks;
    for (int i = 0; i < nat; ++i)
      for (int j = 0; j <= i; ++j) {
        int a = molecule[i].atomNumber, b = molecule[j].atomNumber;
        int r = (a == 1 && b == 1) ? 1 : (a > 1 && b > 1 ? 10 : 1);
        int c = (a == 1 && b == 1) ? 1 : 10;
        for (int k = 0; k < r; ++k)
          for (int l = 0; l < c; ++l) blocks.push_back(ints[i][j][k][l]);
      }
    put("pair_blocks", blocks.data(), (long long)blocks.size());
  }

  Hcore hcore(parameters, infoAOs, ints);
  hcore.Alloc4Matrix("hcoreMatrix");
  hcore.ComputeMatrix();
  auto t2 = std::chrono::high_resolution_clock::now();
  put("hcore", hcore.matrixHold_, (long long)hcore.array1DSize_);

  // Overlap values for every AO pair on different atoms (packed, 0 on same atom).
  {
    Overlap ov;
    std::vector<double> s(N * (N + 1) / 2, 0.0);
    for (size_t i = 0; i < N; ++i)
      for (size_t j = 0; j <= i; ++j)
        if (infoAOs.orbital[i].indexAtom != infoAOs.orbital[j].indexAtom)
          s[i * (i + 1) / 2 + j] = ov.ComputeOverlap(infoAOs.orbital[i], infoAOs.orbital[j]);
    put("overlap_offdiag", s.data(), (long long)s.size());
  }

  double ecc = CoreCoreRepulsion::ComputeRepulsion(parameters, infoAOs, ints);
  put1("e_cc", ecc);

  // Shipped driver at its hard-coded 5e-3 eV threshold.
  double shippedE = 0.0;
  int shippedConv = 0;
  {
    SCFCalculation scf(infoAOs, ints, hcore);
    scf.AllocSCFData();
    scf.ComputeSCF();
    shippedE = scf.finalEnergy;
    shippedConv = scf.energyConverge ? 1 : 0;
    scf.DeallocSCFData();
  }
  put1("shipped_e_el", shippedE);
  put1("shipped_converged", shippedConv);

  // Threshold-parameterised loop (same sequence as ComputeSCF).
  auto t3 = std::chrono::high_resolution_clock::now();
  SCFCalculation scf(infoAOs, ints, hcore);
  scf.AllocSCFData();
  scf.InitialGuess();
  put("guess_p", scf.Pmatrix->matrixHold_, (long long)scf.Pmatrix->array1DSize_);
  put("guess_f", scf.Fmatrix->matrixHold_, (long long)scf.Fmatrix->array1DSize_);
  double eOld = 0.0;
  double e = ElectronicEnergy::ComputeEnergy(N, hcore, *scf.Fmatrix, *scf.Pmatrix);
  std::vector<double> energies(1, e);
  unsigned int steps = 0;
  bool conv = false;
  int lap = 0;
  std::vector<double> firstP, firstF;
  double fockSeconds = 0.0;
  while (!conv && maxSteps > steps) {
    lap = scf.GetEigenValVecOfFmatrix();
    if (lap != 0) break;
    scf.Pmatrix->ComputeMatrix();
    auto f0 = std::chrono::high_resolution_clock::now();
    scf.Fmatrix->ComputeMatrix();
    auto f1 = std::chrono::high_resolution_clock::now();
    fockSeconds += std::chrono::duration<double>(f1 - f0).count();
    if (steps == 0) {
      firstP.assign(scf.Pmatrix->matrixHold_, scf.Pmatrix->matrixHold_ + scf.Pmatrix->array1DSize_);
      firstF.assign(scf.Fmatrix->matrixHold_, scf.Fmatrix->matrixHold_ + scf.Fmatrix->array1DSize_);
    }
    eOld = e;
    e = ElectronicEnergy::ComputeEnergy(N, hcore, *scf.Fmatrix, *scf.Pmatrix);
    energies.push_back(e);
    conv = sameReal(eOld, e, threshold);
    steps++;
  }
  auto t4 = std::chrono::high_resolution_clock::now();
  std::cout.rdbuf(old);

  put("iter1_p", firstP.data(), (long long)firstP.size());
  put("iter1_f", firstF.data(), (long long)firstF.size());
  put("energies", energies.data(), (long long)energies.size());
  put("eigenvalues", scf.eigenVal, (long long)N);
  put("final_p", scf.Pmatrix->matrixHold_, (long long)scf.Pmatrix->array1DSize_);
  put("final_f", scf.Fmatrix->matrixHold_, (long long)scf.Fmatrix->array1DSize_);
  put1("e_el", e);
  put1("e_tot", e + ecc);
  put1("steps", steps);
  put1("converged", conv ? 1 : 0);
  put1("lapack_info", lap);
  put1("threshold", threshold);
  put1("t_integrals_s", std::chrono::duration<double>(t1 - t0).count());
  put1("t_hcore_s", std::chrono::duration<double>(t2 - t1).count());
  put1("t_scf_s", std::chrono::duration<double>(t4 - t3).count());
  put1("t_fock_per_step_s", steps ? fockSeconds / steps : 0.0);
  fclose(g_out);
  printf("%s N=%zu steps=%u E_el=%.10f E_cc=%.10f E_tot=%.10f\n", xyz.c_str(), N, steps, e, ecc,
         e + ecc);
  return 0;
}
