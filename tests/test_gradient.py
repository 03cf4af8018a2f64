"""Nuclear gradients: analytic (sg_gradient) and by central differences (sg_gradient_fd).

Pinned by the mathematics and the published MNDO geometries: a gradient of
a translation- and rotation-invariant energy sums to zero over the atoms and
has no torque; the MNDO equilibrium geometries of Dewar & Thiel (J. Am. Chem.
Soc. 99, 4899 (1977): H2O r_OH 0.943 A, HOH 106.8 deg; CH4 r_CH 1.104 A) are
stationary to the rounding of the published values, and a stretched O-H bond
pulls back.  The CPU tests hold the oracle's own central differences to these;
the GPU tests hold the device gradients to the oracle's.  The analytic
gradient (derivatives of the pair terms at fixed density) is held to
Richardson-extrapolated central differences of the SCF energy, on the device
and in the oracle, which remove the O(step^2) truncation.
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


def formaldehyde_axial():
    """H2CO with the C=O bond on the z axis and the hydrogens in the xz plane:
    the C-O pair takes the local frame's axial case."""
    return [(6, 0.0, 0.0, 0.0), (8, 0.0, 0.0, 1.216), (1, 0.94, 0.0, -0.55),
            (1, -0.94, 0.0, -0.55)]


def water_axial():
    """H2O with one O-H bond on the z axis (an axial X-H pair), slightly
    bent off its equilibrium."""
    return [(8, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 0.97), (1, 0.93, 0.0, -0.25)]


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
        assert np.abs(g - go).max() <= 2e-6
        assert np.abs(g.sum(axis=0)).max() <= 1e-5
        assert np.abs(torque(mol, g)).max() <= 1e-4
    g1 = engine.gradient_fd(water(), step=1e-3)[0]
    g0 = engine.gradient_fd(water(), step=1e-3, diis=0)[0]
    assert np.abs(g1 - g0).max() <= 2e-6


def richardson(fd, mol, h):
    """(4 g(h/2) - g(h)) / 3: central differences with the h^2 term removed."""
    return (4.0 * fd(mol, h / 2) - fd(mol, h)) / 3.0


def test_oracle_richardson_gradient_is_stationary_free_of_truncation():
    """The extrapolated differences of the oracle agree with their own value at
    half the step to 1e-7 eV/A, the accuracy the analytic test relies on."""
    mol = stretched_water()
    f = lambda m, h: oracle_fd(m, step=h, threshold=1e-12)
    g1 = richardson(f, mol, 2e-3)
    g2 = richardson(f, mol, 1e-3)
    assert np.abs(g1 - g2).max() <= 1e-7


def test_oracle_fixed_density_gradient_equals_scf_differences():
    """The premise of the analytic gradient, on the CPU: at a self-consistent
    density (oracle SCF run 40 steps with no energy test) the derivative of
    the energy at fixed density equals the derivative of the SCF energy
    (Richardson differences of full SCFs to 1e-12 eV) within 1e-7 eV/A, for a
    geometry off equilibrium and one with an axial pair."""
    import pyoracle as po
    for mol in (stretched_water(), water_axial()):
        r = po.Molecule(mol).scf(threshold=0.0, max_steps=40, mode="nddo")
        gf = oracle_fixed_density_gradient(mol, r["p"])
        gs = richardson(lambda m, h: oracle_fd(m, step=h, threshold=1e-12), mol, 2e-3)
        assert np.abs(gf - gs).max() <= 1e-7


ANALYTIC_CASES = [("water", water), ("methane", methane), ("stretched water", stretched_water),
                  ("water axial", water_axial), ("formaldehyde axial", formaldehyde_axial),
                  ("pyridine", geometry.pyridine), ("propane", lambda: geometry.alkane(3)),
                  ("(H2O)4", lambda: geometry.water_cluster(4, geometry.SplitMix64(4)))]


def oracle_fixed_density_gradient(mol, p_packed, h=1e-3):
    """Central differences (Richardson-extrapolated) of the oracle's total
    energy at a FIXED density: E(X; P) = E_el(H(X), F(X; P), P) + E_cc(X), the
    pair blocks, Hcore and Fock rebuilt at each displaced geometry.  The
    truncation is O(h^4), negligible at h = 1e-3 A; the rounding of the energy
    sums (~1e-11 eV) enters divided by h, so h = 1e-3 and 2e-3 agree to 1e-8
    eV/A on pyridine and (H2O)4, while h = 1e-4 is off by 1e-7."""
    import pyoracle as po

    def e(m):
        o = po.Molecule(m)
        bl = o.pair_blocks()
        hh = o.hcore(bl)
        f = o.fock(bl, hh, p_packed)
        return po.energy(o.n_ao, hh, f, p_packed) + o.core_core(bl)

    g = np.zeros((len(mol), 3))
    for a in range(len(mol)):
        for k in range(3):
            v = []
            for s in (1.0, -1.0, 0.5, -0.5):
                m = [list(x) for x in mol]
                m[a][1 + k] += s * h
                v.append(e([tuple(x) for x in m]))
            g1 = (v[0] - v[1]) / (2 * h)
            g2 = (v[2] - v[3]) / h
            g[a, k] = (4.0 * g2 - g1) / 3.0
    return g


@pytest.mark.gpu
def test_analytic_gradient_fixed_density_against_oracle(engine):
    """sg_gradient against the oracle's derivative of the same energy
    expression at the device's own converged density (fixed-density central
    differences, so no SCF residual enters either side): within 1e-7 eV/A
    per component, and translation invariance (exact at any fixed density;
    rotation invariance needs self-consistency and is checked on the polished
    gradient below); axial X-X and X-H pairs included."""
    from semigpu_amd import api
    for name, geom in ANALYTIC_CASES + [("ethane", lambda: geometry.alkane(2))]:
        mol = geom()
        b = engine.batch([mol])
        b.set_diis(6)
        r = b.scf(threshold=1e-10, max_steps=200)[0]
        g = b.gradient()
        p = b.packed(0, api.SG_DENSITY)
        b.close()
        assert r["converged"], name
        go = oracle_fixed_density_gradient(mol, p)
        err = np.abs(g - go).max()
        print("analytic", name, "vs oracle at fixed density", err,
              "sum", np.abs(g.sum(axis=0)).max(), flush=True)
        assert err <= 1e-7, (name, np.c_[g, go])
        assert np.abs(g.sum(axis=0)).max() <= 1e-8, name


@pytest.mark.gpu
def test_analytic_gradient_against_differences(engine):
    """The gradient of a polished SCF (Engine.gradient) against
    Richardson-extrapolated central differences of the SCF energy, on the
    device (every case) and in the oracle (the small ones): within 1e-6 eV/A
    per component; no net force or torque."""
    dev_fd = lambda m, h: engine.gradient_fd(m, step=h, threshold=1e-12)[0]
    for name, geom in ANALYTIC_CASES:
        mol = geom()
        g, e, conv = engine.gradient(mol)
        assert conv, name
        gd = richardson(dev_fd, mol, 2e-3)
        print("analytic", name, "vs device differences", np.abs(g - gd).max(), "torque",
              np.abs(torque(mol, g)).max(), flush=True)
        assert np.abs(g - gd).max() <= 1e-6, name
        assert np.abs(g.sum(axis=0)).max() <= 1e-8, name
        assert np.abs(torque(mol, g)).max() <= 1e-7, name
        if len(mol) <= 6:
            go = richardson(lambda m, h: oracle_fd(m, step=h, threshold=1e-12), mol, 2e-3)
            print("analytic", name, "vs oracle differences", np.abs(g - go).max(), flush=True)
            assert np.abs(g - go).max() <= 1e-6, name


@pytest.mark.gpu
def test_analytic_gradient_batch_equals_alone(engine):
    """Gradients of a batch equal each molecule's alone, bit for bit."""
    mols = [geom() for _, geom in ANALYTIC_CASES]
    b = engine.batch(mols)
    b.set_diis(6)
    b.scf(threshold=1e-10, max_steps=200)
    gb = b.gradient()
    b.close()
    i = 0
    for mol in mols:
        g = engine.gradient(mol, threshold=1e-10, polish=0)[0]
        assert np.array_equal(g, gb[i:i + len(mol)])
        i += len(mol)


@pytest.mark.gpu
def test_geometry_optimisation_reaches_published_mndo_geometry(engine):
    """BFGS on the device energy and analytic gradient from distorted starts
    lands on the published MNDO equilibrium (and with the central-difference
    gradient on the same water geometry within 1e-3 A): H2O r_OH 0.943 A, HOH 106.8 deg, dH_f
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
    rf = optimize(engine, start, gradient="fd")
    print("water opt (central differences)", rf["evaluations"], flush=True)
    assert np.abs(np.array([a[1:4] for a in rf["atoms"]]) -
                  np.array([a[1:4] for a in r["atoms"]])).max() <= 1e-3
    s = 1.15 / math.sqrt(3)
    start = [(6, 0.0, 0.0, 0.0), (1, s, s, s), (1, -s, -s, s), (1, -s, s, -s), (1, s, -s, -s * 1.02)]
    r = optimize(engine, start)
    c = np.array(r["atoms"][0][1:4])
    d = [np.linalg.norm(np.array(a[1:4]) - c) for a in r["atoms"][1:]]
    dhf = api.heat_of_formation(r["atoms"], r["e_tot"])
    print("methane opt", r["evaluations"], d, dhf, flush=True)
    assert all(abs(x - 1.104) <= 0.002 for x in d)
    assert abs(dhf + 11.9) <= 0.1
