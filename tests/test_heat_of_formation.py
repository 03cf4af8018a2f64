"""Heat of formation (sg_heat_of_formation), the first step after the SCF.

dH_f = (E_tot - sum E_isol) x 23.060529 + sum dH_f(atom).  Pinned by the
published MNDO values at the MNDO equilibrium geometries (Dewar & Thiel,
J. Am. Chem. Soc. 99, 4899 and 4907 (1977)): H2O -60.9 kcal/mol
(r_OH 0.943 A, HOH 106.8 deg), CH4 -11.9 kcal/mol (r_CH 1.104 A).  The CPU
test takes E_tot from the oracle, the GPU test from the device SCF.
"""
import math

import pytest

from semigpu_amd import api


def water():
    t = math.radians(106.8) / 2
    return [(8, 0.0, 0.0, 0.0), (1, 0.943 * math.sin(t), 0.943 * math.cos(t), 0.0),
            (1, -0.943 * math.sin(t), 0.943 * math.cos(t), 0.0)]


def methane():
    s = 1.104 / math.sqrt(3)
    return [(6, 0.0, 0.0, 0.0), (1, s, s, s), (1, -s, -s, s), (1, -s, s, -s), (1, s, -s, -s)]


PUBLISHED = [("water", water, -60.9), ("methane", methane, -11.9)]


@pytest.mark.parametrize("name,geom,dhf", PUBLISHED)
def test_published_mndo_heats(name, geom, dhf):
    import pyoracle as po
    mol = geom()
    r = po.Molecule(mol).scf(threshold=1e-8, max_steps=200, mode="nddo")
    assert r["converged"]
    assert abs(api.heat_of_formation(mol, r["e_tot"]) - dhf) <= 0.1, name


def test_atoms_at_their_isolated_energy():
    # E_tot equal to the sum of the isolated-atom energies leaves the atoms' heats
    mol = [(8, 0, 0, 0), (6, 1, 0, 0), (7, 2, 0, 0), (1, 3, 0, 0)]
    e_isol = -317.868506 - 120.500606 - 202.566201 - 11.9062760
    assert abs(api.heat_of_formation(mol, e_isol) - (59.559 + 170.89 + 113.0 + 52.102)) < 1e-9
    # one eV below that is 23.060529 kcal/mol lower
    assert abs(api.heat_of_formation(mol, e_isol - 1.0)
               - (59.559 + 170.89 + 113.0 + 52.102 - 23.060529)) < 1e-9


def test_unsupported_element():
    with pytest.raises(api.SemiGpuError, match="Z=9"):
        api.heat_of_formation([(9, 0, 0, 0)], -100.0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,geom,dhf", PUBLISHED)
def test_published_mndo_heats_gpu(engine, name, geom, dhf):
    mol = geom()
    r = engine.batch([mol]).scf(threshold=1e-8, max_steps=200)[0]
    assert r["converged"]
    assert abs(api.heat_of_formation(mol, r["e_tot"]) - dhf) <= 0.1, name
