"""The numpy DIIS restatement (tests/diis_ref.py) over the oracle's Fock
builds, on the CPU: with depth 1 (no extrapolation) it is the oracle's own
SCF (same steps, same energy), and with depth 6 it reaches the oracle SCF's
energy in fewer steps.  The GPU test holds the library's DIIS to it."""
import pytest

import diis_ref
import pyoracle as po
from semigpu_amd import geometry

MOLS = {
    "pyridine": geometry.pyridine,
    "water10": lambda: geometry.water_cluster(10, geometry.SplitMix64(5)),
    "hexane": lambda: geometry.alkane(6),
}


@pytest.mark.parametrize("name", sorted(MOLS))
def test_depth1_is_the_oracle_scf(name):
    mol = MOLS[name]()
    o = po.Molecule(mol).scf(threshold=1e-8, mode="nddo")
    e, steps, conv, _ = diis_ref.diis_scf(mol, 1, 1e-8)
    assert conv and steps == o["steps"]
    assert abs(e - o["e_tot"]) <= 1e-9


@pytest.mark.parametrize("name", sorted(MOLS))
def test_diis_reaches_the_oracle_energy(name):
    mol = MOLS[name]()
    o = po.Molecule(mol).scf(threshold=1e-8, mode="nddo")
    e, steps, conv, _ = diis_ref.diis_scf(mol, 6, 1e-8)
    assert conv and steps < o["steps"]
    assert abs(e - o["e_tot"]) <= 1e-7
