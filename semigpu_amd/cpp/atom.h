// atom.h -- Atom (reference: src/common/atom.h:12-52).
#ifndef SEMIGPU_ATOM_H
#define SEMIGPU_ATOM_H

#include <string>
using std::string;
#include <vector>
using std::vector;

class Atom {
 public:
  Atom();
  int atomNumber;
  string atomSymbol;
  double atomCoordinates[3];  // Angstrom
  double atomWeight;
  int atomValenceElectrons;
  bool statusData;
  void setCoordinates(double, double, double);
  void setCoordinates(vector<double>);
  void setCoordinates(double (&xyz_coordinates)[3]);
  double getXCoordinate();
  double getYCoordinate();
  double getZCoordinate();
  // Symbol or atomic number (1..108); an invalid one clears statusData.
  void setAtomSymbol(string);
  void setAtomNumber(int);
  static void PrintGeometry(const vector<Atom>& molecule);

 protected:
  double xPosition;
  double yPosition;
  double zPosition;

 private:
  void assign(int number);
};

#endif  // SEMIGPU_ATOM_H
