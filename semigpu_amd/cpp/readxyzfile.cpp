// ReadXYZFile (readxyzfile.cpp:27-125 of the reference).
#include "readxyzfile.h"

#include <cctype>
#include <fstream>
#include <iostream>
#include <sstream>

ReadXYZFile::ReadXYZFile() : Natoms(0), open_without_problems(true) {}

bool ReadXYZFile::GetValuesFromFile(string filename, vector<Atom>& molecule) {
  std::ifstream f(filename.c_str());
  if (!f.good()) {
    std::cout << "Error: File " << filename << " could not be opened...\n";
    open_without_problems = false;
    return false;
  }
  if (!GetValuesFromStream(f, molecule)) {
    if (Natoms == 0)
      std::cout << "Error: File " << filename << " does not have the XYZ format..." << std::endl;
    return false;
  }
  return true;
}

bool ReadXYZFile::GetValuesFromStream(std::istream& in, vector<Atom>& molecule) {
  string line;
  open_without_problems = true;
  Natoms = 0;
  while (std::getline(in, line)) {  // skip blank lines between records
    std::istringstream ls(line);
    if (ls >> Natoms) break;
    if (line.find_first_not_of(" \t\r") != string::npos) break;
  }
  if (Natoms <= 0) {
    Natoms = 0;
    open_without_problems = false;
    return false;
  }
  std::getline(in, line);  // comment line
  molecule.assign(Natoms, Atom());
  for (int i = 0; i < Natoms; ++i) {
    string tok;
    vector<double> xyz(3, 0.0);
    if (!(in >> tok >> xyz[0] >> xyz[1] >> xyz[2])) {
      open_without_problems = false;
      return false;
    }
    const bool number = std::isdigit(static_cast<unsigned char>(tok[0])) != 0;
    if (number)
      molecule[i].setAtomNumber(std::atoi(tok.c_str()));
    else
      molecule[i].setAtomSymbol(tok);
    molecule[i].setCoordinates(xyz);
  }
  std::getline(in, line);  // rest of the last atom line
  for (const Atom& a : molecule)
    if (!a.statusData) open_without_problems = false;
  return open_without_problems;
}

bool ReadXYZRecords(const string& filename, vector<vector<Atom>>& molecules) {
  std::ifstream f(filename.c_str());
  if (!f.good()) return false;
  molecules.clear();
  ReadXYZFile r;
  for (;;) {
    vector<Atom> m;
    if (!r.GetValuesFromStream(f, m)) return r.Natoms == 0 && f.eof() && !molecules.empty();
    molecules.push_back(m);
  }
}
