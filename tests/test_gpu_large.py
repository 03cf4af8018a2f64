"""Multi-workgroup eigensolver: molecules too large for one workgroup.

The same molecules go through both eigensolvers (an engine whose large-basis
threshold is lowered to 128 basis functions routes them to the
multi-workgroup one), so the new path is held to the one-workgroup path,
which is itself held to the reference.  Molecules above the one-workgroup
limit (N > 1280) are checked through what fixes the result: eigenpairs
against numpy (residual, orthonormality), the density's idempotency and
trace, the integrals, Hcore, Fock and energy of the converged density against
the oracle's, and the SCF energy against the oracle's SCF of the same cluster,
written once to tests/golden/water_250_oracle.json by
tests/golden/make_oracle_large.py.
"""
import json
import os
import sys

import numpy as np
import pytest

from semigpu_amd import api, geometry

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def large_engine():
    e = api.Engine()
    e.set_large_basis(128)
    yield e
    e.close()


def packed_to_full(p, n):
    a = np.zeros((n, n))
    r, c = np.tril_indices(n)
    a[r, c] = p
    a[c, r] = p
    return a


def fock_of_guess(eng, mol):
    b = eng.batch([mol])
    b.two_center()
    b.hcore()
    b.guess_density()
    b.fock()
    return b


def exact_energy(P, H, F, n):
    """1/2 sum_ij P_ij (H_ij + F_ij) over the full symmetric matrices, with
    exactly rounded summation (math.fsum), from packed lower triangles."""
    import math
    t = P * (H + F)
    diag = np.array([i * (i + 1) // 2 + i for i in range(n)])
    off = np.ones(t.size, dtype=bool)
    off[diag] = False
    return 0.5 * math.fsum(np.concatenate([t[diag], 2.0 * t[off]]).tolist())


def check_eigh(b, tag):
    n = b.n_ao(0)
    F = packed_to_full(b.packed(0, api.SG_FOCK), n)
    b.eigh()
    w, C = b.eigen(0)
    wr = np.linalg.eigh(F)[0]
    scale = np.abs(wr).max()
    ew = np.abs(w - wr).max() / scale
    eo = np.abs(C.T @ C - np.eye(n)).max()
    er = np.abs(F @ C - C * w).max() / scale
    say(tag, "N", n, "eig", ew, "orth", eo, "resid", er)
    assert ew <= 1e-12
    assert eo <= 1e-12
    assert er <= 1e-11


def test_threshold_bounds(engine):
    with pytest.raises(api.SemiGpuError, match="below 128"):
        engine.set_large_basis(64)


@pytest.mark.parametrize("n_water", [40, 70])
def test_eigh_large_path_against_numpy(large_engine, n_water):
    """N = 240 and 420 on the multi-workgroup solver: every eigenpair."""
    say("case eigh large-path waters", n_water)
    mol = geometry.water_cluster(n_water, geometry.SplitMix64(5 + n_water))
    b = fock_of_guess(large_engine, mol)
    check_eigh(b, "large-path")
    b.close()


def test_scf_large_path_equals_one_workgroup(engine, large_engine):
    """SCF energies of the two solvers agree to rounding; molecules below the
    threshold (pyridine, N = 29) keep the one-workgroup path bit for bit."""
    mols = [geometry.water_cluster(40, geometry.SplitMix64(1)), geometry.alkane(40),
            geometry.pyridine(), geometry.water_cluster(70, geometry.SplitMix64(2))]
    say("case scf both solvers", [geometry.n_basis(m) for m in mols])
    ref = engine.batch(mols).scf(threshold=1e-6)
    big = large_engine.batch(mols).scf(threshold=1e-6)
    for m, a, c in zip(mols, ref, big):
        say("N", geometry.n_basis(m), a["steps"], c["steps"], c["e_tot"] - a["e_tot"])
        assert c["converged"] and c["status"] == 0
        assert c["steps"] == a["steps"]
        assert abs(c["e_tot"] - a["e_tot"]) <= 1e-7
    assert big[2] == ref[2]


def test_large_path_batch_equals_alone(large_engine):
    """Large and small molecules in one batch give their results alone, bit for bit."""
    mols = geometry.water_batch(6, seed=3, n_min=2, n_max=10)
    mols.insert(2, geometry.water_cluster(30, geometry.SplitMix64(9)))   # N = 180
    mols.append(geometry.water_cluster(25, geometry.SplitMix64(10)))     # N = 150
    say("case mixed sizes", [geometry.n_basis(m) for m in mols])
    res = large_engine.batch(mols).scf(threshold=1e-6)
    for i in (0, 2, len(mols) - 1):
        alone = large_engine.batch([mols[i]]).scf(threshold=1e-6)[0]
        assert alone == res[i], i


def test_diis_large_path(engine, large_engine):
    """DIIS over the multi-workgroup eigensolver (eager iteration, no graph)
    reaches the one-workgroup path's DIIS energy and the plain SCF's, with
    pyridine below the threshold in the same batch."""
    mols = [geometry.water_cluster(40, geometry.SplitMix64(1)), geometry.pyridine()]
    say("case diis both solvers", [geometry.n_basis(m) for m in mols])
    plain = engine.batch(mols).scf(threshold=1e-8)
    small = engine.batch(mols)
    small.set_diis(6)
    ref = small.scf(threshold=1e-8)
    big = large_engine.batch(mols)
    big.set_diis(6)
    res = big.scf(threshold=1e-8)
    for p, a, c in zip(plain, ref, res):
        say("diis steps", a["steps"], c["steps"], "plain", p["steps"], c["e_tot"] - a["e_tot"])
        assert c["converged"] and c["status"] == 0
        assert abs(c["e_tot"] - a["e_tot"]) <= 1e-7
        assert abs(c["e_tot"] - p["e_tot"]) <= 1e-6
        assert c["steps"] < p["steps"]
    small.close()
    big.close()


def test_mixed_policy_large_path(large_engine):
    mol = geometry.water_cluster(50, geometry.SplitMix64(12))
    say("case mixed policy large path N", geometry.n_basis(mol))
    b = large_engine.batch([mol])
    a = b.scf(threshold=1e-6, policy="fp64")[0]
    c = b.scf(threshold=1e-6, policy="mixed")[0]
    say("fp64", a["e_tot"], "mixed", c["e_tot"])
    assert c["converged"] and abs(c["e_tot"] - a["e_tot"]) <= 1e-4
    b.close()


def test_above_one_workgroup_limit(engine):
    """N = 1500 (250 waters): eigenpairs, density, each SCF ingredient against
    the oracle, and the SCF energy against the oracle's SCF.

    E_el (-2.9e6 eV) and E_cc (+2.9e6 eV) cancel to E_tot = -8.8e4 eV, so two
    FP64 evaluations that sum in different orders differ by ~1e-12 |E_el|;
    the energies are held to that (3e-6 eV here), the matrices to 1e-10."""
    import pyoracle as po
    mol = geometry.water_cluster(250, geometry.SplitMix64(31))
    say("case N=1500 eigh")
    b = fock_of_guess(engine, mol)
    n = b.n_ao(0)
    assert n == 1500
    check_eigh(b, "N=1500")
    b.close()
    say("case N=1500 scf")
    b = engine.batch([mol])
    r = b.scf(threshold=1e-6)[0]
    P = b.packed(0, api.SG_DENSITY)
    H = b.packed(0, api.SG_HCORE)
    F = b.packed(0, api.SG_FOCK)
    Pf = packed_to_full(P, n)
    n_occ = b.info[0][2]
    idem = np.abs(Pf @ Pf - 2.0 * Pf).max()
    say("N=1500 steps", r["steps"], "e_el", r["e_el"], "e_cc", r["e_cc"], "idempotency", idem)
    assert r["converged"] and r["status"] == 0
    assert abs(np.trace(Pf) - 2 * n_occ) <= 1e-9
    assert idem <= 1e-9
    om = po.Molecule(mol)
    blocks = om.pair_blocks()
    assert np.array_equal(b.pair_blocks(0), blocks)
    h_or = om.hcore(blocks)
    assert np.abs(H - h_or).max() <= 1e-10 * np.abs(h_or).max()
    f_or = om.fock(blocks, h_or, P)
    assert np.abs(F - f_or).max() <= 1e-10 * np.abs(f_or).max()
    e_exact = exact_energy(P, H, F, n)
    e_el_or = po.energy(n, H, F, P)
    tol = 1e-12 * abs(r["e_el"])
    say("E_el - exact sum", r["e_el"] - e_exact, "oracle sum - exact", e_el_or - e_exact,
        "tol", tol)
    assert abs(r["e_el"] - e_exact) <= tol
    assert abs(r["e_cc"] - om.core_core(blocks)) <= 1e-12 * abs(r["e_cc"])
    with open(os.path.join(HERE, "golden", "water_250_oracle.json")) as fh:
        g = json.load(fh)
    say("E_tot vs oracle SCF", r["e_tot"] - g["e_tot"])
    assert r["steps"] == g["steps"]
    assert abs(r["e_tot"] - g["e_tot"]) <= tol
    b.close()


def test_water_box_1000(engine):
    """The 1000-water box (3000 atoms, N = 6000) on the multi-workgroup path:
    FP64 SCF to 1e-6 eV, the converged density idempotent with the right
    trace (checked through products with random vectors), the pair blocks bit
    for bit against the oracle's, its Fock matrix against the oracle's from
    the same integrals and density, E_el against the
    exactly rounded sum of its matrices, and the mixed policy within 1e-4 eV
    of FP64."""
    import pyoracle as po
    mol = geometry.water_box()
    b = engine.batch([mol])
    n = b.n_ao(0)
    assert n == 6000
    say("case water box N=6000 fp64 scf")
    r = b.scf(threshold=1e-6, max_steps=60, policy="fp64")[0]
    say("box fp64", r)
    assert r["converged"] and r["status"] == 0
    P = b.packed(0, api.SG_DENSITY)
    H = b.packed(0, api.SG_HCORE)
    F = b.packed(0, api.SG_FOCK)
    Pf = packed_to_full(P, n)
    x = np.random.default_rng(0).standard_normal((n, 4))
    px = Pf @ x
    idem = np.abs(Pf @ px - 2.0 * px).max() / np.abs(px).max()
    say("box idempotency", idem, "trace", np.trace(Pf) - 2 * b.info[0][2])
    assert idem <= 1e-10
    assert abs(np.trace(Pf) - 2 * b.info[0][2]) <= 1e-8
    del Pf
    say("case water box oracle ingredients")
    om = po.Molecule(mol)
    blocks = b.pair_blocks(0)
    # the 4.5 M atom pairs' integrals bit for bit against the oracle's
    assert np.array_equal(blocks, om.pair_blocks())
    f_or = om.fock(blocks, H, P)
    assert np.abs(F - f_or).max() <= 1e-10 * np.abs(f_or).max()
    # E_el = 1/2 sum P (H + F) of the device's own matrices, summed exactly
    e_exact = exact_energy(P, H, F, n)
    e_or = po.energy(n, H, F, P)
    say("box E_el - exact", r["e_el"] - e_exact, "oracle sum - exact", e_or - e_exact)
    assert abs(r["e_el"] - e_exact) <= 1e-12 * abs(e_exact)
    say("case water box mixed scf")
    c = b.scf(threshold=1e-6, max_steps=60, policy="mixed")[0]
    say("box mixed", c, "dE", c["e_tot"] - r["e_tot"])
    assert c["converged"] and abs(c["e_tot"] - r["e_tot"]) <= 1e-4
    say("case water box fp64 scf with DIIS")
    b.set_diis(6)
    d = b.scf(threshold=1e-6, max_steps=60, policy="fp64")[0]
    say("box diis", d, "dE", d["e_tot"] - r["e_tot"])
    assert d["converged"] and d["steps"] < r["steps"]
    assert abs(d["e_tot"] - r["e_tot"]) <= 1e-6
    b.close()
