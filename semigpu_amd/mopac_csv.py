"""MOPAC-comparable CSV dumps of a molecule's SCF ingredients and results.

The reference checks its integrals, Hcore, Fock matrix and energies against
CSV files that src/test/fullTester.bsh:93-172 cuts out of a MOPAC run
(`<name>_twoInt.csv`, `_Hcore.csv`, `_1Fmatrix.csv`, `_Fmatrix.csv`,
`_Pmatrix.csv`, `_energies.csv`), and its test drivers read them with
FileUtils::ReadCSV (src/common/fileutils.cpp:18-45: one row per line, fields
split at commas, reading stops at the first empty line).  This module writes
the same files from this library's results, so that they can be compared
with such MOPAC files, or given to the reference's test drivers in their
place:

  _twoInt.csv    the pair blocks in the reference's order (atom i, atom
                 j <= i, block [k][l]: 1, 10 or 100 values), ten per line
                 (test_twoCenterIntegral.cc:123-180 reads them as one list);
  _Hcore.csv, _1Fmatrix.csv, _Fmatrix.csv, _Pmatrix.csv
                 N lines of N fields "%.6f", the lower triangle with zeros
                 above it (make_matrix of fullTester.bsh: the lower triangle
                 MOPAC prints, empty fields set to 0); the Fock matrix of the
                 first SCF iteration (built from the guess density), the
                 converged one and the converged density;
  _energies.csv  TOTAL_ENERGY, ELECTRONIC_ENERGY, CORE-CORE_REPULSION, each
                 "name,value,EV" (print_energies of fullTester.bsh).
"""
from __future__ import annotations

import os
from typing import Dict, List, Sequence

import numpy as np

from . import api

SUFFIXES = ("twoInt", "Hcore", "1Fmatrix", "Fmatrix", "Pmatrix", "energies")


def read_csv(path: str) -> List[List[str]]:
    """Rows of fields as FileUtils::ReadCSV returns them."""
    rows = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line:
                break
            rows.append(line.split(","))
    return rows


def write_matrix_csv(path: str, packed: np.ndarray, n: int) -> None:
    """Packed lower triangle (row-major, r(r+1)/2 + c) as N x N fields."""
    packed = np.asarray(packed, dtype=np.float64)
    if packed.size != n * (n + 1) // 2:
        raise ValueError("packed size %d does not match N = %d" % (packed.size, n))
    with open(path, "w") as f:
        for i in range(n):
            row = packed[i * (i + 1) // 2: i * (i + 1) // 2 + i + 1]
            fields = ["%.6f" % v for v in row] + ["%.6f" % 0.0] * (n - i - 1)
            f.write(",".join(fields) + "\n")


def write_two_int_csv(path: str, blocks: np.ndarray) -> None:
    """Flattened pair blocks, ten values per line."""
    blocks = np.asarray(blocks, dtype=np.float64)
    with open(path, "w") as f:
        for s in range(0, blocks.size, 10):
            f.write(",".join("%.6f" % v for v in blocks[s:s + 10]) + "\n")


def write_energies_csv(path: str, e_tot: float, e_el: float, e_cc: float) -> None:
    with open(path, "w") as f:
        f.write("TOTAL_ENERGY,%.8f,EV\n" % e_tot)
        f.write("ELECTRONIC_ENERGY,%.8f,EV\n" % e_el)
        f.write("CORE-CORE_REPULSION,%.8f,EV\n" % e_cc)


def dump_molecule(engine: api.Engine, atoms: Sequence, prefix: str, threshold: float = 5e-3,
                  max_steps: int = 100) -> Dict[str, str]:
    """Runs the molecule (atoms: (Z, x, y, z) in Angstrom) through the stages
    and the SCF on the GPU and writes the six files `<prefix>_<suffix>.csv`.
    Returns {suffix: path} and the SCF result under "result"."""
    b = engine.batch([atoms])
    try:
        n = b.n_ao(0)
        b.two_center()
        b.hcore()
        b.core_core()
        b.guess_density()
        b.fock()
        paths = {s: "%s_%s.csv" % (prefix, s) for s in SUFFIXES}
        d = os.path.dirname(prefix)
        if d:
            os.makedirs(d, exist_ok=True)
        write_two_int_csv(paths["twoInt"], b.pair_blocks(0))
        write_matrix_csv(paths["Hcore"], b.packed(0, api.SG_HCORE), n)
        write_matrix_csv(paths["1Fmatrix"], b.packed(0, api.SG_FOCK), n)
        r = b.scf(threshold=threshold, max_steps=max_steps)[0]
        write_matrix_csv(paths["Fmatrix"], b.packed(0, api.SG_FOCK), n)
        write_matrix_csv(paths["Pmatrix"], b.packed(0, api.SG_DENSITY), n)
        write_energies_csv(paths["energies"], r["e_tot"], r["e_el"], r["e_cc"])
    finally:
        b.close()
    out: Dict[str, object] = dict(paths)
    out["result"] = r
    return out
