"""Writes tests/golden/water_box_1000_oracle.json: the oracle's FP64 SCF of the
1000-water box (geometry.water_box(), N = 6000), with the energy of every
iteration (E_0 from the guess Fock matrix, then E_1 .. E_steps), the
reference for the per-iteration error of the device's policies on this box.

The compiled reference needs days at this size (its Fock build is O(N^4)),
so the fixture comes from the oracle (oracle/mndo_oracle.c, pair-blocked Fock
and its own dense eigensolver, held to the reference's results on the other
fixtures of this folder).  About two hours on one core.

Usage: python tests/golden/make_oracle_box.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from semigpu_amd import geometry  # noqa: E402
import pyoracle  # noqa: E402


def main():
    mol = geometry.water_box()
    t = time.time()
    o = pyoracle.Molecule(mol).scf(threshold=1e-6, max_steps=60, mode="nddo")
    rec = {"molecule": "geometry.water_box() (1000 waters, seed 1000)",
           "n_ao": geometry.n_basis(mol), "threshold_eV": 1e-6,
           "e_el": o["e_el"], "e_cc": o["e_cc"], "e_tot": o["e_tot"], "steps": o["steps"],
           "converged": o["converged"], "energies_el": [float(x) for x in o["energies"]],
           "oracle_seconds": round(time.time() - t, 1)}
    with open(os.path.join(ROOT, "tests", "golden", "water_box_1000_oracle.json"), "w") as fh:
        json.dump(rec, fh, indent=1)
    print(rec)


if __name__ == "__main__":
    main()
