// Atom (atom.cpp:14-138 of the reference): element bookkeeping of one atom.
#include "atom.h"

#include <iomanip>
#include <iostream>

namespace {

const char* const kSymbols[108] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",
    "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn",
    "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh",
    "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd",
    "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re",
    "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
    "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db",
    "Sg", "Bh", "Hs"};

// Standard atomic weights (IUPAC, abridged) of the first three periods; the
// SCF path does not use the weight, so heavier elements carry 0.
const double kWeights[18] = {1.008,  4.0026, 6.94,   9.0122, 10.81,  12.011,
                             14.007, 15.999, 18.998, 20.180, 22.990, 24.305,
                             26.982, 28.085, 30.974, 32.06,  35.45,  39.95};

// Valence electrons = MNDO core charge for the first three periods.
int valence(int n) {
  if (n <= 2) return n;
  if (n <= 10) return n - 2;
  if (n <= 18) return n - 10;
  return -1;
}

}  // namespace

Atom::Atom()
    : atomNumber(0), atomCoordinates{0.0, 0.0, 0.0}, atomWeight(0.0), atomValenceElectrons(0),
      statusData(true), xPosition(0.0), yPosition(0.0), zPosition(0.0) {}

void Atom::setCoordinates(double x, double y, double z) {
  atomCoordinates[0] = xPosition = x;
  atomCoordinates[1] = yPosition = y;
  atomCoordinates[2] = zPosition = z;
}

void Atom::setCoordinates(vector<double> v) { setCoordinates(v.at(0), v.at(1), v.at(2)); }

void Atom::setCoordinates(double (&xyz)[3]) { setCoordinates(xyz[0], xyz[1], xyz[2]); }

double Atom::getXCoordinate() { return xPosition; }
double Atom::getYCoordinate() { return yPosition; }
double Atom::getZCoordinate() { return zPosition; }

void Atom::assign(int number) {
  atomNumber = number;
  atomWeight = number <= 18 ? kWeights[number - 1] : 0.0;
  atomValenceElectrons = valence(number);
  if (atomValenceElectrons < 0) statusData = false;
}

void Atom::setAtomSymbol(string element) {
  for (int i = 0; i < 108; ++i)
    if (element == kSymbols[i]) {
      atomSymbol = element;
      assign(i + 1);
      return;
    }
  atomNumber = 0;
  atomSymbol = element + " Invalid atomic symbol";
  statusData = false;
}

void Atom::setAtomNumber(int number) {
  if (number > 0 && number <= 108) {
    atomSymbol = kSymbols[number - 1];
    assign(number);
    return;
  }
  atomNumber = number;
  atomSymbol = " Invalid atomic number";
  statusData = false;
}

void Atom::PrintGeometry(const vector<Atom>& molecule) {
  std::cout << "Geometry :" << std::endl;
  for (const Atom& a : molecule)
    std::cout << std::setw(3) << a.atomSymbol << std::setw(10) << a.atomCoordinates[0]
              << std::setw(10) << a.atomCoordinates[1] << std::setw(10) << a.atomCoordinates[2]
              << std::endl;
}
