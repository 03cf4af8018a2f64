// readxyzfile.h -- XYZ reader (reference: src/common/readxyzfile.h:26-55).
#ifndef SEMIGPU_READXYZFILE_H
#define SEMIGPU_READXYZFILE_H

#include <istream>
#include <string>
using std::string;
#include <vector>
using std::vector;

#include "atom.h"

class ReadXYZFile {
 public:
  ReadXYZFile();
  int Natoms;
  bool open_without_problems;
  // Line 1: atom count; line 2: comment; then one "Z-or-symbol x y z" (Angstrom)
  // per atom.  False when the file cannot be opened, the count is 0 or an atom
  // is invalid.
  bool GetValuesFromFile(string, vector<Atom>&);
  // The same from an open stream positioned at the count line; reads exactly
  // one molecule, so concatenated XYZ records are read one call at a time.
  bool GetValuesFromStream(std::istream&, vector<Atom>&);
};

// Every molecule of a file of concatenated XYZ records (one multi-molecule
// request file); false at the first malformed record.
bool ReadXYZRecords(const string& filename, vector<vector<Atom>>& molecules);

#endif  // SEMIGPU_READXYZFILE_H
