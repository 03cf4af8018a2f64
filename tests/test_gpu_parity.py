"""Parity of the HIP path (through the C ABI) with the reference CPU path.

Bars (also stated in DESIGN.md):
  pair blocks            bit-identical to the reference
  Hcore, Fock            |gpu - ref| <= 1e-10 * max|ref|  (normwise relative)
  energies               |gpu - ref| <= 1e-6 eV at SCF threshold 1e-6 eV, same step count
  batch vs alone         bit-identical (no floating-point atomics, fixed reduction order;
                         mixed_eig32 batches above 512 molecules: the SCF threshold)
Reference values are the committed fixtures written by the compiled reference
(tests/golden); sizes beyond the fixtures are checked against the oracle.
"""
import sys

import numpy as np
import pytest

from conftest import atoms_of, golden_names, load_golden
from semigpu_amd import api, geometry

pytestmark = pytest.mark.gpu

REL = 1e-10
FULL = golden_names(full_only=True)
LARGE = [n for n in golden_names() if n not in FULL]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def normwise(a, ref):
    return float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("name", FULL)
def test_integrals_hcore_fock(engine, name):
    say("case", name)
    g = load_golden(name)
    b = engine.batch([atoms_of(g)])
    b.two_center()
    assert np.array_equal(b.pair_blocks(0), g["pair_blocks"])
    b.hcore()
    assert normwise(b.packed(0, api.SG_HCORE), g["hcore"]) <= REL
    b.guess_density()
    b.fock()
    assert normwise(b.packed(0, api.SG_FOCK), g["guess_f"]) <= REL
    b.set_packed(0, api.SG_DENSITY, g["iter1_p"])
    b.fock()
    assert normwise(b.packed(0, api.SG_FOCK), g["iter1_f"]) <= REL
    assert abs(b.core_core()[0] - float(g["e_cc"])) <= 1e-9
    b.close()


@pytest.mark.parametrize("name", LARGE)
def test_large_sampled(engine, name):
    say("case", name)
    g = load_golden(name)
    b = engine.batch([atoms_of(g)])
    b.two_center()
    blk = b.pair_blocks(0)
    assert np.array_equal(blk[g["pair_blocks_idx"]], g["pair_blocks_val"])
    b.hcore()
    h = b.packed(0, api.SG_HCORE)
    assert np.abs(h[g["hcore_idx"]] - g["hcore_val"]).max() <= REL * np.abs(g["hcore_val"]).max()
    r = b.scf(threshold=float(g["threshold"]), max_steps=100)[0]
    assert r["steps"] == int(g["steps"]) and r["converged"]
    assert abs(r["e_el"] - float(g["e_el"])) <= 1e-6
    assert abs(r["e_tot"] - float(g["e_tot"])) <= 1e-6
    b.close()


def test_scf_energies_batch(engine):
    names = FULL
    say("batch of", len(names))
    mols = [atoms_of(load_golden(n)) for n in names]
    b = engine.batch(mols)
    res = b.scf(threshold=1e-6, max_steps=100)
    for n, r in zip(names, res):
        g = load_golden(n)
        assert r["converged"] == bool(g["converged"]), n
        assert r["steps"] == int(g["steps"]), n
        assert abs(r["e_el"] - float(g["e_el"])) <= 1e-6, n
        assert abs(r["e_cc"] - float(g["e_cc"])) <= 1e-9, n
        assert abs(r["e_tot"] - float(g["e_tot"])) <= 1e-6, n
    b.close()


def test_shipped_threshold(engine):
    """Default threshold 5e-3 eV: the reference's own ComputeSCF result."""
    names = ["ref_04_pyridine", "ref_03_formaldehyde", "ref_02_benzene", "ref_02_styrene"]
    b = engine.batch([atoms_of(load_golden(n)) for n in names])
    res = b.scf(threshold=5e-3, max_steps=100)
    for n, r in zip(names, res):
        assert abs(r["e_el"] - float(load_golden(n)["shipped_e_el"])) <= 1e-6, n
    b.close()


def test_batch_equals_alone_bitwise(engine):
    mols = geometry.water_batch(12, seed=77)
    mols.insert(3, geometry.alkane(7))
    mols.insert(7, geometry.pyridine())
    big = engine.batch(mols).scf(threshold=1e-6)
    for i in (0, 3, 7, len(mols) - 1):
        alone = engine.batch([mols[i]]).scf(threshold=1e-6)[0]
        say("alone", i, alone["e_tot"], big[i]["e_tot"])
        assert alone == big[i]


@pytest.mark.parametrize("policy", ["fp64", "mixed"])
def test_two_groups_equal_alone_bitwise(engine, policy):
    """A batch with molecules on both sides of 256 basis functions runs its
    large and small molecules' eigensolve, density and Fock build on two
    streams at once (graph capture and direct launches): every molecule's
    result equals the molecule alone, bit for bit."""
    rng = geometry.SplitMix64(31)
    mols = [geometry.water_cluster(n, rng) for n in (60, 5, 47, 30, 64, 2, 44, 12)]
    mols.insert(2, geometry.alkane(45))  # N = 272
    mols.insert(5, geometry.pyridine())
    sizes = [geometry.n_basis(m) for m in mols]
    assert any(s > 256 for s in sizes) and any(s <= 256 for s in sizes)
    for graphs in (True, False):
        engine.graphs(graphs)
        try:
            big = engine.batch(mols).scf(threshold=1e-6, policy=policy)
            for i, m in enumerate(mols):
                alone = engine.batch([m]).scf(threshold=1e-6, policy=policy)[0]
                assert alone == big[i], (graphs, i, sizes[i])
        finally:
            engine.graphs(True)
    # the same batch as one group on one stream
    engine.overlap(False)
    try:
        assert engine.batch(mols).scf(threshold=1e-6, policy=policy) == big
    finally:
        engine.overlap(True)


def test_eigensolver_properties(engine):
    g = load_golden("ref_02_benzene")
    n = int(g["n_ao"])
    b = engine.batch([atoms_of(g)])
    b.two_center()
    b.hcore()
    b.set_packed(0, api.SG_DENSITY, g["final_p"])
    b.fock()
    b.eigh()
    w, C = b.eigen(0)
    import pyoracle as po
    F = po.packed_to_full(b.packed(0, api.SG_FOCK), n)
    wr = np.linalg.eigh(F)[0]
    scale = np.abs(wr).max()
    assert np.abs(w - wr).max() <= 1e-12 * scale
    assert np.abs(C.T @ C - np.eye(n)).max() <= 1e-12
    assert np.abs(F @ C - C * w).max() <= 1e-11 * scale
    b.close()


def _fused_range_molecules():
    """Molecules above the on-chip reduction's block (256 rows), whose first
    N - 256 columns k_tridiag reduces: N = 257 (an OH group beside 42
    waters), 258, 260, 300 and 384, and N = 386, which puts the batch's head
    launch on the form for N > 384."""
    oh = geometry.water_cluster(42, geometry.SplitMix64(3))
    oh = oh + [(8, 40.0, 40.0, 40.0), (1, 40.0, 40.0, 40.97)]
    return [oh, geometry.water_cluster(43, geometry.SplitMix64(4)), geometry.alkane(43),
            geometry.water_cluster(50, geometry.SplitMix64(6)),
            geometry.water_cluster(64, geometry.SplitMix64(8)), geometry.alkane(64)]


def _check_eigenpairs(b, i, n):
    import pyoracle as po
    w, C = b.eigen(i)
    F = po.packed_to_full(b.packed(i, api.SG_FOCK), n)
    wr = np.linalg.eigh(F)[0]
    scale = np.abs(wr).max()
    ew = np.abs(w - wr).max() / scale
    eo = np.abs(C.T @ C - np.eye(n)).max()
    er = np.abs(F @ C - C * w).max() / scale
    say("N", n, "eig", ew, "orth", eo, "resid", er)
    assert ew <= 1e-12 and eo <= 1e-12 and er <= 1e-11
    return w, C


def test_eigensolver_head_range(engine):
    """Eigenpairs against numpy across N = 257 .. 386 (1 to 130 columns
    before the on-chip block), each molecule alone and all in one batch (bit
    for bit the same, although the batch's head launch takes the form for
    N > 384)."""
    mols = _fused_range_molecules()
    ns = [geometry.n_basis(m) for m in mols]
    say("case eigh head range", ns)
    assert ns == [257, 258, 260, 300, 384, 386]
    b = engine.batch(mols)
    b.two_center()
    b.hcore()
    b.guess_density()
    b.fock()
    b.eigh()
    for i, (m, n) in enumerate(zip(mols, ns)):
        w, C = _check_eigenpairs(b, i, n)
        a = engine.batch([m])
        a.two_center()
        a.hcore()
        a.guess_density()
        a.fock()
        a.eigh()
        wa, Ca = a.eigen(0)
        assert np.array_equal(w, wa) and np.array_equal(C, Ca)
        a.close()
    b.close()


def test_scf_head_range_against_oracle(engine):
    """FP64 SCF of (H2O)45 (N = 270) and C43H88 (N = 260) against the oracle's
    SCF: step counts and energies."""
    import pyoracle as po
    mols = [geometry.water_cluster(45, geometry.SplitMix64(21)), geometry.alkane(43)]
    say("case scf head range", [geometry.n_basis(m) for m in mols])
    res = engine.batch(mols).scf(threshold=1e-6)
    for mol, r in zip(mols, res):
        o = po.Molecule(mol).scf(threshold=1e-6, mode="nddo")
        say("N", geometry.n_basis(mol), r["steps"], o["steps"], r["e_tot"] - o["e_tot"])
        assert r["converged"] and r["steps"] == o["steps"]
        assert abs(r["e_tot"] - o["e_tot"]) <= 1e-6


def test_water_clusters_against_oracle(engine):
    import pyoracle as po
    mols = geometry.water_batch(6, seed=5, n_min=2, n_max=10)
    res = engine.batch(mols).scf(threshold=1e-6)
    for mol, r in zip(mols, res):
        o = po.Molecule(mol).scf(threshold=1e-6, mode="nddo")
        assert r["steps"] == o["steps"]
        assert abs(r["e_tot"] - o["e_tot"]) <= 1e-6


@pytest.mark.parametrize("waters", [64, 200])
def test_converged_density_idempotent(engine, waters):
    """FP64 SCF of N = 384 and N = 1200 (one workgroup's eigensolver, one
    inverse-iteration solve per vector): the converged density is idempotent
    (P^2 = 2P for the closed shell) with the right trace."""
    import pyoracle as po
    mol = geometry.water_cluster(waters, geometry.SplitMix64(19))
    b = engine.batch([mol])
    r = b.scf(threshold=1e-6)[0]
    n = b.n_ao(0)
    P = po.packed_to_full(b.packed(0, api.SG_DENSITY), n)
    n_occ = b.info[0][2]
    b.close()
    idem = np.abs(P @ P - 2.0 * P).max()
    say("N", n, "steps", r["steps"], "idempotency", idem, "trace", np.trace(P) - 2 * n_occ)
    assert r["converged"]
    assert abs(np.trace(P) - 2 * n_occ) <= 1e-9
    assert idem <= 1e-9


def test_energy_history_against_oracle(engine):
    """The per-iteration record (sg_scf_record_energies) against the oracle's
    SCF energies step by step (scfcalculation.cpp:66-146: E_0 from the guess
    Fock matrix, then one energy per step): same number of steps and every
    E_el(k) within 1e-6 eV, in one batch; under the mixed policy the record
    has one entry per step and its last FP64 steps meet the oracle's
    converged energy within 1e-4 eV."""
    import pyoracle
    mols = [geometry.pyridine(), geometry.water_cluster(8, geometry.SplitMix64(8)),
            geometry.alkane(10)]
    b = engine.batch(mols)
    b.record_energies(101)
    res = b.scf(threshold=1e-6)
    hist = [b.energy_history(i) for i in range(len(mols))]
    resm = b.scf(threshold=1e-6, policy="mixed")
    histm = [b.energy_history(i) for i in range(len(mols))]
    b.close()
    for i, mol in enumerate(mols):
        o = pyoracle.Molecule(mol).scf(threshold=1e-6, mode="nddo")
        eo = np.asarray(o["energies"])
        d = np.abs(hist[i] - eo).max() if len(hist[i]) == len(eo) else np.inf
        say("history", i, len(hist[i]), len(eo), d)
        assert res[i]["steps"] == o["steps"] and len(hist[i]) == o["steps"] + 1
        assert d <= 1e-6
        assert hist[i][-1] == res[i]["e_el"]
        assert len(histm[i]) == resm[i]["steps"] + 1
        assert abs(histm[i][-1] - o["e_el"]) <= 1e-4


def test_unsupported_element_rejected(engine):
    with pytest.raises(api.SemiGpuError, match="Z=9"):
        engine.batch([[(9, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 0.9)]])


def test_step_cap_reports_not_converged(engine):
    b = engine.batch([geometry.pyridine()])
    r = b.scf(threshold=1e-12, max_steps=2)[0]
    assert r["steps"] == 2 and not r["converged"] and r["status"] == 0
    r0 = engine.batch([geometry.pyridine()]).scf(threshold=1e-6, max_steps=0)[0]
    assert r0["steps"] == 0 and not r0["converged"]


MIXED = ["mixed", "mixed_bf16", "mixed_eig32"]


@pytest.mark.parametrize("policy", MIXED)
def test_mixed_policy_within_1e4(engine, policy):
    """Reduced-precision steps until |dE| < 1e-3 eV, then FP64:
    |E - E_ref| <= 1e-4 eV on every fixture."""
    names = FULL + LARGE
    mols = [atoms_of(load_golden(n)) for n in names]
    res = engine.batch(mols).scf(threshold=1e-6, max_steps=100, policy=policy)
    for n, r in zip(names, res):
        g = load_golden(n)
        say(policy, n, r["steps"], r["e_tot"] - float(g["e_tot"]))
        assert r["converged"], n
        assert abs(r["e_tot"] - float(g["e_tot"])) <= 1e-4, n


@pytest.mark.parametrize("policy", MIXED)
def test_mixed_policy_water_batch(engine, policy):
    mols = geometry.water_batch(64, seed=11)
    b = engine.batch(mols)
    ref = b.scf(threshold=1e-6, policy="fp64")
    assert not b.fp32_steps().any()
    mix = b.scf(threshold=1e-6, policy=policy)
    s32 = b.fp32_steps()
    worst = max(abs(a["e_tot"] - c["e_tot"]) for a, c in zip(ref, mix))
    say("mixed vs fp64 worst |dE|", worst, "FP32 steps", s32.min(), s32.max())
    assert all(r["converged"] for r in mix)
    assert worst <= 1e-4
    # every molecule starts in FP32 and ends with two FP64 steps
    assert all(1 <= n32 <= r["steps"] - 2 for n32, r in zip(s32, mix))


@pytest.mark.parametrize("policy", MIXED)
def test_mixed_policy_large_batch_vs_alone(engine, policy):
    """A 600-molecule batch against its molecules alone, under every mixed
    policy: every kernel is per molecule in a fixed order (mixed_eig32 reduces
    its FP32 and FP64 molecules in separate launches at every batch size), so
    the results are equal bit for bit, and within 1e-4 eV of FP64."""
    mols = geometry.water_batch(600, seed=5, n_min=2, n_max=10)
    b = engine.batch(mols)
    big = b.scf(threshold=1e-6, policy=policy)
    assert all(r["converged"] for r in big)
    ref = b.scf(threshold=1e-6, policy="fp64")
    for i in (0, 1, 299, 599):
        alone = engine.batch([mols[i]]).scf(threshold=1e-6, policy=policy)[0]
        say(policy, "alone", i, alone["steps"], alone["e_tot"] - big[i]["e_tot"],
            alone["e_tot"] - ref[i]["e_tot"])
        assert alone["converged"]
        assert alone == big[i], i
        assert abs(alone["e_tot"] - ref[i]["e_tot"]) <= 1e-4, i
    b.close()


def test_mixed_policy_split_launches_above_384(engine):
    """mixed_eig32: the FP32 reduction for N > 384 (eight wavefronts, one workgroup per CU):
    300 molecules, one with N = 420, take separate FP32 and FP64 launches;
    alone, the N = 420 molecule takes the combined one.  Both within the SCF
    threshold of each other and within 1e-4 eV of FP64."""
    mols = geometry.water_batch(299, seed=8, n_min=2, n_max=6)
    mols.insert(150, geometry.water_cluster(70, geometry.SplitMix64(70)))  # N = 420
    b = engine.batch(mols)
    mix = b.scf(threshold=1e-6, policy="mixed_eig32")
    s32 = b.fp32_steps()
    ref = b.scf(threshold=1e-6, policy="fp64")
    assert all(r["converged"] for r in mix)
    assert max(abs(a["e_tot"] - c["e_tot"]) for a, c in zip(ref, mix)) <= 1e-4
    assert s32[150] >= 1
    alone = engine.batch([mols[150]]).scf(threshold=1e-6, policy="mixed_eig32")[0]
    say("N=420 mixed batch", mix[150]["steps"], "alone", alone["steps"],
        alone["e_tot"] - mix[150]["e_tot"], "vs fp64", mix[150]["e_tot"] - ref[150]["e_tot"])
    assert alone["converged"] and abs(alone["e_tot"] - mix[150]["e_tot"]) <= 1e-6
    b.close()


def test_fp32_eigensolver_step_density(engine):
    """The first SCF step of the mixed_eig32 policy diagonalises the FP64 guess
    Fock matrix on the FP32 path (reduction, bisection to 1e-10 ||T||, one solve,
    back-transform and density in FP32).  Its density is held to the FP64
    step's: |dP| <= 1e-4, trace 2 n_occ to 1e-4, and P P = 2 P to 1e-4."""
    mols = [geometry.water_cluster(64, geometry.SplitMix64(4)),   # N = 384
            geometry.water_cluster(100, geometry.SplitMix64(5)),  # N = 600 (eight wavefronts)
            geometry.alkane(12), geometry.pyridine()]
    b = engine.batch(mols)
    b.scf(threshold=1e-6, max_steps=1, policy="mixed_eig32")
    assert all(n32 == 1 for n32 in b.fp32_steps())
    p32 = [b.packed(i, api.SG_DENSITY) for i in range(len(mols))]
    b.scf(threshold=1e-6, max_steps=1, policy="fp64")
    for i, m in enumerate(mols):
        n = b.n_ao(i)
        p64 = b.packed(i, api.SG_DENSITY)
        P = np.zeros((n, n))
        r, c = np.tril_indices(n)
        P[r, c] = p32[i]
        P[c, r] = p32[i]
        dp = np.abs(p32[i] - p64).max()
        idem = np.abs(P @ P - 2.0 * P).max()
        say("fp32 step N", n, "dP", dp, "idempotency", idem, "trace", np.trace(P) - 2 * b.info[i][2])
        assert dp <= 1e-4
        assert abs(np.trace(P) - 2 * b.info[i][2]) <= 1e-4
        assert idem <= 1e-4
    b.close()


@pytest.mark.parametrize("policy,bar", [("mixed", 1e-5), ("mixed_bf16", 5e-4)])
def test_reduced_density_product_first_step(engine, policy, bar):
    """The first SCF step of the mixed and mixed_bf16 policies diagonalises
    the FP64 guess Fock matrix with the FP64 eigensolver and forms the density
    product in FP32, or as split BF16 (hi hi + hi lo + lo hi, FP32
    accumulation), on the matrix cores.  Its density against the FP64 step's:
    |dP| within the bar (FP32 rounding: about 1e-7 per product; split BF16
    about 2^-16), the trace 2 n_occ and P P = 2 P to the same bar."""
    mols = [geometry.water_cluster(64, geometry.SplitMix64(4)),   # N = 384
            geometry.water_cluster(100, geometry.SplitMix64(5)),  # N = 600
            geometry.alkane(12), geometry.pyridine()]
    b = engine.batch(mols)
    b.scf(threshold=1e-6, max_steps=1, policy=policy)
    assert all(n32 == 1 for n32 in b.fp32_steps())
    pr = [b.packed(i, api.SG_DENSITY) for i in range(len(mols))]
    b.scf(threshold=1e-6, max_steps=1, policy="fp64")
    for i, m in enumerate(mols):
        n = b.n_ao(i)
        p64 = b.packed(i, api.SG_DENSITY)
        P = np.zeros((n, n))
        r, c = np.tril_indices(n)
        P[r, c] = pr[i]
        P[c, r] = pr[i]
        dp = np.abs(pr[i] - p64).max()
        idem = np.abs(P @ P - 2.0 * P).max()
        say(policy, "first step N", n, "dP", dp, "idempotency", idem,
            "trace", np.trace(P) - 2 * b.info[i][2])
        assert dp <= bar
        assert abs(np.trace(P) - 2 * b.info[i][2]) <= bar * n
        assert idem <= 4 * bar
    b.close()


def test_damped_scf(engine):
    """Opt-in density damping reaches the undamped SCF's energy; batch equals
    alone and graph replay equals eager under damping; damping 0 is the
    reference's SCF again, bit for bit."""
    mols = [geometry.pyridine(), geometry.water_cluster(20, geometry.SplitMix64(20)),
            geometry.alkane(10)]
    ref = engine.batch(mols).scf(threshold=1e-6)
    b = engine.batch(mols)
    b.set_damping(0.3)
    damp = b.scf(threshold=1e-6)
    for r, d in zip(ref, damp):
        say("damped", d["steps"], "undamped", r["steps"], d["e_tot"] - r["e_tot"])
        assert d["converged"] and abs(d["e_tot"] - r["e_tot"]) <= 1e-5
    alone = engine.batch([mols[0]])
    alone.set_damping(0.3)
    assert alone.scf(threshold=1e-6)[0] == damp[0]
    alone.close()
    engine.timing(True)
    eager = b.scf(threshold=1e-6)
    engine.timing(False)
    engine.timing_reset()
    assert eager == damp
    with pytest.raises(api.SemiGpuError, match="damping"):
        b.set_damping(1.0)
    b.set_damping(0.0)
    assert b.scf(threshold=1e-6) == ref
    b.close()


def test_diis_scf(engine):
    """Opt-in Pulay DIIS reaches the plain SCF's energy in no more steps;
    batch equals alone and graph replay equals eager under DIIS; a converged
    molecule's Fock matrix is the one of its final density; depth 0 is the
    reference's SCF again, bit for bit."""
    mols = [geometry.pyridine(), geometry.water_cluster(20, geometry.SplitMix64(20)),
            geometry.alkane(10), geometry.water_cluster(60, geometry.SplitMix64(61))]
    ref = engine.batch(mols).scf(threshold=1e-8)
    b = engine.batch(mols)
    b.set_diis(6)
    dis = b.scf(threshold=1e-8)
    for r, d in zip(ref, dis):
        say("diis", d["steps"], "plain", r["steps"], d["e_tot"] - r["e_tot"])
        assert d["converged"] and abs(d["e_tot"] - r["e_tot"]) <= 1e-6
        assert d["steps"] <= r["steps"]
    assert sum(d["steps"] for d in dis) < sum(r["steps"] for r in ref)
    f_scf = b.packed(3, api.SG_FOCK)
    b.fock()
    np.testing.assert_allclose(b.packed(3, api.SG_FOCK), f_scf, rtol=0, atol=1e-12)
    alone = engine.batch([mols[1]])
    alone.set_diis(6)
    assert alone.scf(threshold=1e-8)[0] == dis[1]
    alone.close()
    engine.timing(True)
    eager = b.scf(threshold=1e-8)
    engine.timing(False)
    engine.timing_reset()
    assert eager == dis
    mixed = b.scf(threshold=1e-8, policy="mixed")
    for r, d in zip(ref, mixed):
        say("diis mixed", d["steps"], d["e_tot"] - r["e_tot"])
        assert d["converged"] and abs(d["e_tot"] - r["e_tot"]) <= 1e-6
    for bad in (1, 9, -1):
        with pytest.raises(api.SemiGpuError, match="DIIS"):
            b.set_diis(bad)
    b.set_diis(3)
    for r, d in zip(ref, b.scf(threshold=1e-8)):
        assert d["converged"] and abs(d["e_tot"] - r["e_tot"]) <= 1e-6
    b.set_diis(0)
    assert b.scf(threshold=1e-8) == ref
    b.close()


def test_graph_replay_equals_eager(engine):
    """The captured SCF iteration (replayed once per step, with eager setup
    work and host synchronisation between replays) gives the eager results."""
    mols = geometry.water_batch(24, seed=99, n_min=2, n_max=30)
    b = engine.batch(mols)
    for policy in ("fp64", "mixed"):
        g1 = b.scf(threshold=1e-6, policy=policy)
        engine.timing(True)
        eager = b.scf(threshold=1e-6, policy=policy)
        engine.timing(False)
        engine.timing_reset()
        g2 = b.scf(threshold=1e-6, policy=policy)
        assert g1 == eager == g2, policy
    b.close()


def test_large_cluster_against_oracle(engine):
    """N = 600 takes the long-row tridiagonalisation path (N > 384)."""
    import pyoracle as po
    mol = geometry.water_cluster(100, geometry.SplitMix64(4242))
    r = engine.batch([mol]).scf(threshold=1e-6)[0]
    o = po.Molecule(mol).scf(threshold=1e-6, mode="nddo")
    say("N=600", r["steps"], o["steps"], r["e_tot"] - o["e_tot"])
    assert r["converged"] and r["steps"] == o["steps"]
    assert abs(r["e_tot"] - o["e_tot"]) <= 1e-6


def test_eigensolver_largest_basis(engine):
    """N = 1200 (the largest cluster of the serving stream) against numpy."""
    mol = geometry.water_cluster(200, geometry.SplitMix64(77))
    b = engine.batch([mol])
    n = b.n_ao(0)
    assert n == 1200
    b.two_center()
    b.hcore()
    b.guess_density()
    b.fock()
    b.eigh()
    w, C = b.eigen(0)
    import pyoracle as po
    F = po.packed_to_full(b.packed(0, api.SG_FOCK), n)
    wr = np.linalg.eigh(F)[0]
    scale = np.abs(wr).max()
    assert np.abs(w - wr).max() <= 1e-12 * scale
    assert np.abs(C.T @ C - np.eye(n)).max() <= 1e-12
    assert np.abs(F @ C - C * w).max() <= 1e-11 * scale
    b.close()


EDGE = [
    [(1, 0.0, 0.0, 0.0)],                       # N = 1, no occupied orbital
    [(6, 0.0, 0.0, 0.0)], [(7, 0.0, 0.0, 0.0)], [(8, 0.0, 0.0, 0.0)],   # N = 4, one atom
    [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 0.74)],  # N = 2 (no Householder reflector)
    [(1, 0.0, 0.0, 0.0), (8, 0.0, 0.0, 0.97)],  # odd electron count (n_occ = floor)
    [(6, 0.0, 0.0, 0.0), (8, 0.0, 0.0, 1.13)],
]


def test_edge_molecules_against_oracle(engine):
    """Single atoms, H2, OH and CO, alone and in one batch with a large
    cluster: energies and step counts as the oracle's, batch equal to alone."""
    import pyoracle as po
    mols = EDGE + [geometry.water_cluster(40, geometry.SplitMix64(3))]
    say("case edge molecules", [geometry.n_basis(m) for m in mols])
    res = engine.batch(mols).scf(threshold=1e-6)
    for m, r in zip(mols, res):
        o = po.Molecule(m).scf(threshold=1e-6, mode="nddo")
        assert r["status"] == 0 and r["converged"] == o["converged"]
        assert r["steps"] == o["steps"]
        assert abs(r["e_tot"] - o["e_tot"]) <= 1e-6
    for i in range(len(EDGE)):
        assert engine.batch([mols[i]]).scf(threshold=1e-6)[0] == res[i], i


def test_empty_molecule_rejected(engine):
    L = engine.L
    import ctypes
    offs = np.array([0, 0, 1], dtype=np.int32)
    Z = np.array([1], dtype=np.int32)
    xyz = np.zeros(3)
    h = ctypes.c_void_p()
    rc = L.sg_batch_create(engine.h, 2, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           Z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(h))
    assert rc == -1 and b"no atoms" in L.sg_last_error()


def test_diis_against_numpy_restatement(engine):
    """The device DIIS (FP64 policy) takes the steps of the numpy restatement
    over the oracle's Fock builds and lands on its energy."""
    import diis_ref
    mols = [geometry.pyridine(), geometry.water_cluster(10, geometry.SplitMix64(5)),
            geometry.alkane(6), geometry.water_cluster(20, geometry.SplitMix64(20))]
    for depth in (3, 6):
        b = engine.batch(mols)
        b.set_diis(depth)
        res = b.scf(threshold=1e-8)
        b.close()
        for m, r in zip(mols, res):
            e, steps, conv, _ = diis_ref.diis_scf(m, depth, 1e-8)
            say("diis depth", depth, len(m), "gpu", r["steps"], "numpy", steps, r["e_tot"] - e)
            assert r["converged"] and conv
            assert r["steps"] == steps
            assert abs(r["e_tot"] - e) <= 1e-8
