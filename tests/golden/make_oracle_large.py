"""Writes tests/golden/water_250_oracle.json: the oracle's SCF of a 250-water
cluster (N = 1500, above the one-workgroup eigensolver limit).

The compiled reference needs hours at this size, so the fixture comes from the
oracle (oracle/mndo_oracle.c, whose results are held to the reference's on the
other fixtures of this folder).  About two minutes on one core.

Usage: python tests/golden/make_oracle_large.py
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
    mol = geometry.water_cluster(250, geometry.SplitMix64(31))
    t = time.time()
    o = pyoracle.Molecule(mol).scf(threshold=1e-6, mode="nddo")
    rec = {"molecule": "geometry.water_cluster(250, SplitMix64(31))",
           "n_ao": geometry.n_basis(mol), "threshold_eV": 1e-6,
           "e_el": o["e_el"], "e_cc": o["e_cc"], "e_tot": o["e_tot"], "steps": o["steps"],
           "converged": o["converged"], "oracle_seconds": round(time.time() - t, 1)}
    with open(os.path.join(ROOT, "tests", "golden", "water_250_oracle.json"), "w") as fh:
        json.dump(rec, fh, indent=1)
    print(rec)


if __name__ == "__main__":
    main()
