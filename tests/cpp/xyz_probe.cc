// Host-only test driver of the XYZ reader and Atom (no device calls):
// prints "ok Natoms" then one "Z symbol x y z valence status" line per atom,
// or "fail Natoms"; with -r, reads concatenated records and prints their sizes.
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>

#include "readxyzfile.h"

int main(int argc, char* argv[]) {
  if (argc > 2 && !std::strcmp(argv[1], "-r")) {
    vector<vector<Atom>> ms;
    const bool ok = ReadXYZRecords(argv[2], ms);
    std::cout << (ok ? "ok" : "fail") << " " << ms.size();
    for (auto& m : ms) std::cout << " " << m.size();
    std::cout << std::endl;
    return 0;
  }
  ReadXYZFile r;
  vector<Atom> mol;
  const bool ok = r.GetValuesFromFile(argv[1], mol);
  std::cout << (ok ? "ok " : "fail ") << r.Natoms << std::endl;
  std::cout << std::setprecision(17);
  for (Atom& a : mol)
    std::cout << a.atomNumber << " " << a.atomSymbol.substr(0, 2) << " " << a.getXCoordinate()
              << " " << a.getYCoordinate() << " " << a.getZCoordinate() << " "
              << a.atomValenceElectrons << " " << (a.statusData ? 1 : 0) << std::endl;
  return 0;
}
