"""Nuclear gradients by central differences (sg_gradient_fd).

Pinned by the mathematics and the published MNDO geometries: a gradient of
a translation- and rotation-invariant energy sums to zero over the atoms and
has no torque; the MNDO equilibrium geometries of Dewar & Thiel (J. Am. Chem.
Soc. 99, 4899 (1977): H2O r_OH 0.943 A, HOH 106.8 deg; CH4 r_CH 1.104 A) are
stationary to the rounding of the published values, and a stretched O-H bond
pulls back.  The CPU tests hold the oracle's own central differences to these;
the GPU tests hold the device gradient to the oracle's.
"""
import numpy as np
import pytest

from semigpu_amd import geometry
from test_heat_of_formation import methane, water

STEP = 1e-3
THRESH = 1e-9


def stretched_water():
    w = water()
    return [w[0], (1, w[1][1] * 1.05, w[1][2] * 1.05, 0.0), w[2]]


def oracle_fd(mol, step=STEP, threshold=THRESH):
    import pyoracle as po
    g = np.zeros((len(mol), 3))
    for a in range(len(mol)):
        for k in range(3):
            e = []
            for sgn in (1.0, -1.0):
                m = [list(x) for x in mol]
                m[a][1 + k] += sgn * step
                r = po.Molecule([tuple(x) for x in m]).scf(threshold=threshold, max_steps=200,
                                                            mode="nddo")
                assert r["converged"]
                e.append(r["e_tot"])
            g[a, k] = (e[0] - e[1]) / (2 * step)
    return g


def torque(mol, g):
    r = np.array([a[1:4] for a in mol], dtype=np.float64)
    return np.cross(r, g).sum(axis=0)


@pytest.mark.parametrize("geom", [water, methane])
def test_oracle_gradient_stationary_at_published_geometry(geom):
    mol = geom()
    g = oracle_fd(mol)
    assert np.abs(g).max() <= 0.05
    assert np.abs(g.sum(axis=0)).max() <= 1e-5
    assert np.abs(torque(mol, g)).max() <= 1e-4


def test_oracle_gradient_stretched_bond_pulls_back():
    mol = stretched_water()
    g = oracle_fd(mol)
    oh = np.array(mol[1][1:4]) - np.array(mol[0][1:4])
    assert np.abs(g).max() >= 1.0
    assert g[1] @ oh > 0          # dE/dr_OH > 0: the force on H points to O
    assert np.abs(g.sum(axis=0)).max() <= 1e-5


@pytest.mark.gpu
def test_gradient_fd_against_oracle(engine):
    """Device gradient (one batch of 6 n_atoms displaced copies, DIIS) against
    the oracle's central differences of the same step."""
    for mol in (water(), methane(), stretched_water(), geometry.pyridine()):
        g, e, conv = engine.gradient_fd(mol, step=STEP, threshold=THRESH)
        go = oracle_fd(mol)
        r = engine.batch([mol]).scf(threshold=THRESH, max_steps=200)[0]
        print("gradient", len(mol), np.abs(g - go).max(), np.abs(g.sum(axis=0)).max(), flush=True)
        assert conv
        assert abs(e - r["e_tot"]) <= 1e-8
        assert np.abs(g - go).max() <= 1e-5
        assert np.abs(g.sum(axis=0)).max() <= 1e-5
        assert np.abs(torque(mol, g)).max() <= 1e-4
    g1 = engine.gradient_fd(water(), step=1e-3)[0]
    g0 = engine.gradient_fd(water(), step=1e-3, diis=0)[0]
    assert np.abs(g1 - g0).max() <= 1e-5


@pytest.mark.gpu
def test_geometry_optimisation_reaches_published_mndo_geometry(engine):
    """BFGS on the device energy and gradient from distorted starts lands on
    the published MNDO equilibrium: H2O r_OH 0.943 A, HOH 106.8 deg, dH_f
    -60.9 kcal/mol; CH4 r_CH 1.104 A, dH_f -11.9 kcal/mol."""
    import math
    from semigpu_amd import api
    from semigpu_amd.optimize import optimize
    t = math.radians(100.0) / 2
    start = [(8, 0.0, 0.0, 0.0), (1, 1.0 * math.sin(t), 1.0 * math.cos(t), 0.0),
             (1, -0.98 * math.sin(t), 0.98 * math.cos(t), 0.05)]
    r = optimize(engine, start)
    o, h1, h2 = (np.array(a[1:4]) for a in r["atoms"])
    d1, d2 = np.linalg.norm(h1 - o), np.linalg.norm(h2 - o)
    ang = math.degrees(math.acos((h1 - o) @ (h2 - o) / (d1 * d2)))
    dhf = api.heat_of_formation(r["atoms"], r["e_tot"])
    print("water opt", r["evaluations"], d1, d2, ang, dhf, np.abs(r["gradient"]).max(), flush=True)
    assert np.abs(r["gradient"]).max() <= 2e-3
    assert abs(d1 - 0.943) <= 0.002 and abs(d2 - 0.943) <= 0.002
    assert abs(ang - 106.8) <= 0.3
    assert abs(dhf + 60.9) <= 0.1
    s = 1.15 / math.sqrt(3)
    start = [(6, 0.0, 0.0, 0.0), (1, s, s, s), (1, -s, -s, s), (1, -s, s, -s), (1, s, -s, -s * 1.02)]
    r = optimize(engine, start)
    c = np.array(r["atoms"][0][1:4])
    d = [np.linalg.norm(np.array(a[1:4]) - c) for a in r["atoms"][1:]]
    dhf = api.heat_of_formation(r["atoms"], r["e_tot"])
    print("methane opt", r["evaluations"], d, dhf, flush=True)
    assert all(abs(x - 1.104) <= 0.002 for x in d)
    assert abs(dhf + 11.9) <= 0.1
