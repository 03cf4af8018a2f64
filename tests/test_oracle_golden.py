"""Pins the oracle (oracle/mndo_oracle.c) before it is trusted as the checker.

* Bit for bit against fixtures written by the compiled reference
  (tests/golden/make_golden.py -> oracle/_ref/ref_dump): pair blocks, Hcore,
  STO-6G overlaps, guess Fock, first-iteration Fock, core-core energy.
* The pair-blocked Fock against the reference's O(N^4) element loop.
* The eigensolver standing in for dsyev against LAPACK (numpy) and invariants.
* Closed forms: analytic s/p Gaussian overlaps, self-overlap 1, 1/R limit of
  (ss|ss), rigid-motion invariance, idempotent converged density.
* The README's published pyridine energies (README.md:57-61).
"""
import math

import numpy as np
import pytest

from conftest import atoms_of, golden_names, load_golden
import pyoracle as po
from semigpu_amd import geometry

FULL = golden_names(full_only=True)
ALL = golden_names()
SMALL_FOCK = [n for n in FULL if int(load_golden(n)["n_ao"]) <= 40]


@pytest.mark.parametrize("name", FULL)
def test_blocks_hcore_bitexact(name):
    g = load_golden(name)
    m = po.Molecule(atoms_of(g))
    blk = m.pair_blocks()
    assert np.array_equal(blk, g["pair_blocks"]), "pair blocks differ from the reference"
    h = m.hcore(blk)
    assert np.array_equal(h, g["hcore"]), "Hcore differs from the reference"
    assert m.core_core(blk) == float(g["e_cc"])


@pytest.mark.parametrize("name", SMALL_FOCK)
def test_fock_ref_mode_bitexact(name):
    g = load_golden(name)
    m = po.Molecule(atoms_of(g))
    blk = g["pair_blocks"]
    h = g["hcore"]
    f0 = m.fock(blk, h, m.guess_density(), mode="ref")
    assert np.array_equal(f0, g["guess_f"])
    f1 = m.fock(blk, h, g["iter1_p"], mode="ref")
    assert np.array_equal(f1, g["iter1_f"])
    # pair-blocked form of the same sum: summation order only
    f1b = m.fock(blk, h, g["iter1_p"], mode="nddo")
    scale = np.abs(g["iter1_f"]).max()
    assert np.abs(f1b - g["iter1_f"]).max() <= 1e-12 * scale


@pytest.mark.parametrize("name", ["ref_04_pyridine", "ref_02_styrene", "water_08"])
def test_overlaps_bitexact(name):
    g = load_golden(name)
    m = po.Molecule(atoms_of(g))
    ov = g["overlap_offdiag"]
    n = m.n_ao
    for i in range(n):
        for j in range(i + 1):
            k = i * (i + 1) // 2 + j
            if ov[k] != 0.0:
                assert m.overlap(i, j) == ov[k]


@pytest.mark.parametrize("name", [n for n in ALL if n not in FULL])
def test_sampled_large(name):
    """Large fixtures hold a fixed sample of entries (C60, C40H82, (H2O)32, ...)."""
    g = load_golden(name)
    m = po.Molecule(atoms_of(g))
    blk = m.pair_blocks()
    assert np.array_equal(blk[g["pair_blocks_idx"]], g["pair_blocks_val"])
    h = m.hcore(blk)
    assert np.array_equal(h[g["hcore_idx"]], g["hcore_val"])
    assert m.core_core(blk) == float(g["e_cc"])


@pytest.mark.parametrize("name", [n for n in FULL if int(load_golden(n)["n_ao"]) <= 64])
def test_scf_energies(name):
    g = load_golden(name)
    m = po.Molecule(atoms_of(g))
    r = m.scf(threshold=float(g["threshold"]), max_steps=100, mode="nddo")
    assert r["converged"] == bool(g["converged"])
    assert r["steps"] == int(g["steps"])
    assert abs(r["e_el"] - float(g["e_el"])) < 1e-6
    assert abs(r["e_tot"] - float(g["e_tot"])) < 1e-6
    assert np.abs(r["eigenvalues"] - g["eigenvalues"]).max() < 1e-8


def test_shipped_threshold_matches_reference_driver():
    """5e-3 eV threshold reproduces SCFCalculation::ComputeSCF as shipped."""
    for name in ("ref_04_pyridine", "ref_03_formaldehyde", "ref_02_benzene"):
        g = load_golden(name)
        r = po.Molecule(atoms_of(g)).scf(threshold=5e-3, mode="ref")
        assert abs(r["e_el"] - float(g["shipped_e_el"])) < 1e-9


def test_readme_pyridine_numbers():
    g = load_golden("ref_04_pyridine")
    r = po.Molecule(atoms_of(g)).scf(threshold=5e-3, mode="ref")
    # README.md:57-61 prints 6 significant digits
    assert float("%.6g" % r["e_el"]) == -3360.88
    assert float("%.6g" % r["e_cc"]) == 2444.36
    assert float("%.6g" % r["e_tot"]) == -916.525
    # README.md:93-95: formaldehyde -873.482 eV
    g = load_golden("ref_03_formaldehyde")
    r = po.Molecule(atoms_of(g)).scf(threshold=5e-3, mode="ref")
    assert float("%.6g" % r["e_el"]) == -873.482


# --------------------------------------------------------------------------- closed forms
def _prim_1d(la, lb, p, pa, pb):
    base = math.sqrt(math.pi / p)
    if la == 0 and lb == 0:
        return base
    if la == 1 and lb == 0:
        return base * pa
    if la == 0 and lb == 1:
        return base * pb
    return base * (pa * pb + 0.5 / p)


def _sto_overlap_analytic(ea, la, A, eb, lb, B):
    """Obara-Saika s/p overlap of the STO-6G contractions, independent of the oracle."""
    tab = _STO[ea], _STO[eb]
    ca = tab[0]["cS" if la == 0 else "cP"]
    xa = tab[0]["eS" if la == 0 else "eP"]
    cb = tab[1]["cS" if lb == 0 else "cP"]
    xb = tab[1]["eS" if lb == 0 else "eP"]
    ab2 = sum((A[k] - B[k]) ** 2 for k in range(3))
    tot = 0.0
    for i in range(6):
        for j in range(6):
            p = xa[i] + xb[j]
            mu = xa[i] * xb[j] / p
            o = ca[i] * cb[j] * math.exp(-mu * ab2)
            for k in range(3):
                P = (xa[i] * A[k] + xb[j] * B[k]) / p
                o *= _prim_1d(int(la == k + 1), int(lb == k + 1), p, P - A[k], P - B[k])
            tot += o
    return tot


# STO-6G tables restated from the reference's effective values (STO-6G.cpp "Sparrow" blocks)
_STO = {
    1: {"eS": [4.098792193376860e+01, 7.515090619202102e+00, 2.102451537968084e+00,
               7.222488471249888e-01, 2.804703624957924e-01, 1.155131838418812e-01],
        "cS": [1.057956874127020e-01, 1.596794562489045e-01, 2.097264517638662e-01,
               2.069126553471125e-01, 1.144014417359658e-01, 1.840525175585220e-02]},
    6: {"eS": [8.846144274163005e+01, 1.622292919221501e+01, 4.558993095148346e+00,
               6.519461344814428e-01, 2.958932522507792e-01, 1.411098199523801e-01],
        "cS": [-8.534093658582062e-02, -1.190837759635380e-01, -1.145234971659361e-01,
               1.730334274890658e-01, 1.607236215755646e-01, 2.810826497961667e-02],
        "eP": [1.875086664726173e+01, 4.889844711496806e+00, 1.749633028685196e+00,
               7.313800499385337e-01, 3.344367808801327e-01, 1.581099099775126e-01],
        "cP": [4.407314493338078e-01, 5.331737162464377e-01, 5.445171182695406e-01,
               3.904445162839797e-01, 1.454563993779586e-01, 1.494840672526637e-02]},
    8: {"eS": [2.018091738330405e+02, 3.700975064350955e+01, 1.040053837613310e+01,
               1.487300079059505e+00, 6.750282487920551e-01, 3.219171573709536e-01],
        "cS": [-1.584154624760238e-01, -2.210511414259479e-01, -2.125860518260085e-01,
               3.211960348235300e-01, 2.983457630220672e-01, 5.217641116074215e-02],
        "eP": [4.277679392805792e+01, 1.115531796470053e+01, 3.991479056714438e+00,
               1.668514542173559e+00, 7.629585088943974e-01, 3.606998633356254e-01],
        "cP": [1.235684643746845e+00, 1.494866260646997e+00, 1.526669908217376e+00,
               1.094694498739664e+00, 4.078180472887439e-01, 4.191104734371579e-02]},
}


def test_overlap_closed_form():
    mol = [(8, 0.0, 0.0, 0.0), (6, 0.7, -0.4, 1.1), (1, -0.9, 0.3, 0.2)]
    m = po.Molecule(mol)
    bohr = 1.8897259886
    types = []
    for a in mol:
        for t in range(1 if a[0] == 1 else 4):
            types.append((a[0], t, [c * bohr for c in a[1:]]))
    for i in range(m.n_ao):
        for j in range(i):
            ea, la, A = types[i]
            eb, lb, B = types[j]
            ref = _sto_overlap_analytic(ea, la, A, eb, lb, B)
            assert abs(m.overlap(i, j) - ref) <= 1e-13 * max(1.0, abs(ref))


@pytest.mark.parametrize("z", [1, 6, 7, 8])
def test_self_overlap_is_one(z):
    """Normalisation is folded into the STO-6G coefficients: <mu|mu> = 1."""
    m = po.Molecule([(z, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 5.0)])
    for a in range(1 if z == 1 else 4):
        assert abs(m.overlap(a, a) - 1.0) < 1e-9


def test_ss_ss_large_distance_limit():
    for z1, z2 in ((1, 1), (6, 8), (7, 1)):
        R = 150.0
        m = po.Molecule([(z1, 0.0, 0.0, 0.0), (z2, 0.3, -0.2, R)])
        blk = m.pair_blocks()
        # block (1,0) starts after the one-center block of atom 0
        off = 1 if z1 == 1 else 100
        rb = math.sqrt(0.3 ** 2 + 0.2 ** 2 + R ** 2) * 1.8897259886
        assert abs(blk[off] / (27.211386 / rb) - 1.0) < 1e-4


def test_rigid_motion_invariance():
    base = geometry.water_cluster(3, geometry.SplitMix64(5))
    rot = geometry.SplitMix64(11).rotation()
    moved = []
    for z, x, y, w in base:
        v = [sum(rot[r][c] * (x, y, w)[c] for c in range(3)) for r in range(3)]
        moved.append((z, v[0] + 3.0, v[1] - 1.0, v[2] + 0.5))
    e1 = po.Molecule(base).scf(1e-9, mode="nddo")
    e2 = po.Molecule(moved).scf(1e-9, mode="nddo")
    assert abs(e1["e_tot"] - e2["e_tot"]) < 1e-7
    perm = base[3:] + base[:3]
    e3 = po.Molecule(perm).scf(1e-9, mode="nddo")
    assert abs(e1["e_tot"] - e3["e_tot"]) < 1e-7


def test_converged_density_idempotent():
    g = load_golden("ref_04_pyridine")
    m = po.Molecule(atoms_of(g))
    r = m.scf(1e-8, mode="nddo")
    P = po.packed_to_full(r["p"], m.n_ao)
    assert np.abs(P @ P - 2 * P).max() < 1e-10
    assert abs(np.trace(P) - 2 * m.n_occ) < 1e-10


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 150])
def test_eigh_against_lapack(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a + a.T
    a[0, :] *= 3.0
    a[:, 0] *= 3.0
    w, v = po.eigh(a)
    wr = np.linalg.eigh(a)[0]
    assert np.abs(w - wr).max() < 1e-12 * max(1.0, np.abs(wr).max())
    assert np.abs(a @ v - v * w).max() < 1e-11 * max(1.0, np.abs(wr).max())
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12


def test_eigh_degenerate_and_fock():
    """Benzene's Fock matrix has degenerate pi levels: the eigenvectors are
    arbitrary inside each degenerate pair, the density is not."""
    g = load_golden("ref_02_benzene")
    n = int(g["n_ao"])
    F = po.packed_to_full(g["final_f"], n)
    w, v = po.eigh(F)
    wr, vr = np.linalg.eigh(F)
    assert np.abs(w - wr).max() < 1e-12 * np.abs(wr).max()
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12
    assert np.abs(F @ v - v * w).max() < 1e-11 * np.abs(wr).max()
    nocc = po.Molecule(atoms_of(g)).n_occ
    P = po.packed_to_full(po.density(n, nocc, v.T.reshape(-1)), n)
    Pr = 2.0 * vr[:, :nocc] @ vr[:, :nocc].T
    assert np.abs(P - Pr).max() < 1e-11
