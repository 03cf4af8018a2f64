// semiGPU -- MNDO energy of one molecule from an XYZ file (the reference's
// src/App/semiGPU.cc:20-116 command line and output), every stage on the GPU.
//
//   semiGPU [--threshold EV] file.xyz
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <vector>

#include "atom.h"
#include "atomicOrbitals.h"
#include "corecorerepulsion.h"
#include "hcore.h"
#include "MNDO_parameters.h"
#include "readxyzfile.h"
#include "scfcalculation.h"
#include "screenutils.h"
#include "twocenterintegral.h"

using std::cout;
using std::endl;
using std::setw;

static int run(int argc, char* argv[]);

int main(int argc, char* argv[]) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << "semiGPU: error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}

static int run(int argc, char* argv[]) {
  const auto start = std::chrono::high_resolution_clock::now();
  if (argc < 2) {
    cout << "usage: semiGPU [--threshold EV] file.xyz" << endl;
    return EXIT_FAILURE;
  }
  double threshold = 5.0e-3;
  for (int i = 1; i + 1 < argc; ++i)
    if (!std::strcmp(argv[i], "--threshold")) threshold = std::atof(argv[i + 1]);
  const string filename = argv[argc - 1];
  cout << "XYZ file : " << filename << endl;
  ReadXYZFile reader;
  vector<Atom> molecule;
  if (!reader.GetValuesFromFile(filename, molecule)) {
    cout << "Problem to read xyz file " << endl;
    return EXIT_FAILURE;
  }
  cout << "Read XYZ file successful" << endl;
  Atom::PrintGeometry(molecule);
  ScreenUtils::PrintScrStarLine();

  ListAtomicOrbitals infoAOs;
  infoAOs.SetOrbitals(molecule);
  if (!infoAOs.statusData) {
    cout << "Problem to set the atomic orbitals" << endl;
    return EXIT_FAILURE;
  }
  double**** all2CenterIntegrals = nullptr;
  if (!TwoCenterIntegral::Alloc4AllTwoCenterIntegral(molecule, all2CenterIntegrals)) {
    cout << "Problem to alloc all2CenterIntegrals" << endl;
    return EXIT_FAILURE;
  }
  cout << "Correct Alloc: all2CenterIntegrals" << endl;
  MNDOparameter parameters;
  TwoCenterIntegral twoCenInt(parameters);
  twoCenInt.ComputeAllTwoCenterIntegral(infoAOs, all2CenterIntegrals);
  TwoCenterIntegral::To_device(molecule, all2CenterIntegrals);

  Hcore hcore(parameters, infoAOs, all2CenterIntegrals);
  if (hcore.Alloc4Matrix("hcoreMatrix"))
    cout << "Correct Alloc: hcoreMatrix" << endl;
  else
    cout << "Bad Alloc: hcoreMatrix " << endl;
  hcore.ComputeMatrix();
  SCFCalculation SCFprocess(infoAOs, all2CenterIntegrals, hcore);
  SCFprocess.thresholdEnergy = threshold;
  if (!SCFprocess.AllocSCFData()) {
    cout << "Bad Alloc: SCFData " << endl;
    return EXIT_FAILURE;
  }
  cout << "Correct Alloc: SCFData" << endl;
  ScreenUtils::PrintScrStarLine();
  SCFprocess.ComputeSCF();

  const double electronicEnergy = SCFprocess.finalEnergy;
  ScreenUtils::PrintScrStarLine();
  cout << "  Electronic Energy = " << setw(10) << electronicEnergy << " eV" << endl;
  ScreenUtils::PrintScrStarLine();
  const double eCC = CoreCoreRepulsion::ComputeRepulsion(parameters, infoAOs, all2CenterIntegrals);
  cout << "Core-Core Repulsion = " << setw(10) << eCC << " eV" << endl;
  ScreenUtils::PrintScrStarLine();
  cout << "       Total Energy = " << setw(10) << electronicEnergy + eCC << " eV" << endl;
  ScreenUtils::PrintScrStarLine();
  cout << std::setprecision(12) << "  E_el = " << electronicEnergy << "  E_cc = " << eCC
       << "  E_tot = " << electronicEnergy + eCC << "  converged = " << SCFprocess.energyConverge
       << std::setprecision(6) << endl;

  cout << "Dealloc array: SCFData" << endl;
  SCFprocess.DeallocSCFData();
  cout << "Dealloc array: hcoreMatrix" << endl;
  hcore.Dealloc4Matrix();
  cout << "Dealloc array: all2CenterIntegral" << endl;
  TwoCenterIntegral::Dealloc4AllTwoCenterIntegral(molecule, all2CenterIntegrals);
  const auto end = std::chrono::high_resolution_clock::now();
  ScreenUtils::PrintScrStarLine();
  cout << "Time taken for calculation: "
       << std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count() / 1000.0
       << " s" << endl;
  return EXIT_SUCCESS;
}
