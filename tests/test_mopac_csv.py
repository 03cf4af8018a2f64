"""MOPAC-comparable CSV dumps: the layout the reference's test drivers read
(src/test/fullTester.bsh:93-172, FileUtils::ReadCSV), written from the
oracle's matrices on the CPU and from the device's results on the GPU."""
import numpy as np
import pytest

from conftest import atoms_of, load_golden
from semigpu_amd import geometry, mopac_csv


def lower_from_rows(rows, n):
    """What test_hcore.cc / test_fockmatrix.cc take from a matrix CSV."""
    return np.array([float(rows[i][j]) for i in range(n) for j in range(i + 1)])


def test_matrix_layout_round_trip(tmp_path):
    n = 7
    packed = np.random.default_rng(1).standard_normal(n * (n + 1) // 2) * 30.0
    path = str(tmp_path / "m_Hcore.csv")
    mopac_csv.write_matrix_csv(path, packed, n)
    rows = mopac_csv.read_csv(path)
    assert len(rows) == n and all(len(r) == n for r in rows)
    assert np.abs(lower_from_rows(rows, n) - packed).max() <= 5e-7
    assert all(float(rows[i][j]) == 0.0 for i in range(n) for j in range(i + 1, n))
    with pytest.raises(ValueError):
        mopac_csv.write_matrix_csv(path, packed[:-1], n)


def test_two_int_and_energies_layout(tmp_path):
    import pyoracle as po
    mol = geometry.pyridine()
    blocks = po.Molecule(mol).pair_blocks()
    path = str(tmp_path / "p_twoInt.csv")
    mopac_csv.write_two_int_csv(path, blocks)
    rows = mopac_csv.read_csv(path)
    assert all(len(r) == 10 for r in rows[:-1]) and 1 <= len(rows[-1]) <= 10
    flat = np.array([float(v) for r in rows for v in r])
    assert flat.size == blocks.size
    assert np.abs(flat - blocks).max() <= 5e-7
    path = str(tmp_path / "p_energies.csv")
    mopac_csv.write_energies_csv(path, -916.2346415792, -3365.6188171792, 2449.3841756)
    rows = mopac_csv.read_csv(path)
    assert [r[0] for r in rows] == ["TOTAL_ENERGY", "ELECTRONIC_ENERGY", "CORE-CORE_REPULSION"]
    assert all(r[2] == "EV" for r in rows)
    assert float(rows[1][1]) == pytest.approx(-3365.6188171792, abs=1e-8)  # test_scfcalculation.cc:137


@pytest.mark.gpu
def test_dump_pyridine_against_fixtures(engine, tmp_path):
    """The six files of pyridine from the device, read as the reference's test
    drivers read them, against the reference fixtures (6-decimal fields)."""
    g = load_golden("cfg1_pyridine")
    n = int(g["n_ao"])
    out = mopac_csv.dump_molecule(engine, atoms_of(g), str(tmp_path / "pyridine"),
                                  threshold=float(g["threshold"]))
    rows = mopac_csv.read_csv(out["twoInt"])
    flat = np.array([float(v) for r in rows for v in r])
    assert np.abs(flat - g["pair_blocks"]).max() <= 5e-7
    for suffix, key in (("Hcore", "hcore"), ("1Fmatrix", "guess_f"), ("Fmatrix", "final_f"),
                        ("Pmatrix", "final_p")):
        rows = mopac_csv.read_csv(out[suffix])
        assert len(rows) == n and all(len(r) == n for r in rows), suffix
        assert np.abs(lower_from_rows(rows, n) - g[key]).max() <= 1e-6, suffix
    rows = mopac_csv.read_csv(out["energies"])
    assert float(rows[0][1]) == pytest.approx(float(g["e_tot"]), abs=1e-6)
    assert float(rows[1][1]) == pytest.approx(float(g["e_el"]), abs=1e-6)
    assert float(rows[2][1]) == pytest.approx(float(g["e_cc"]), abs=1e-6)
    assert out["result"]["steps"] == int(g["steps"])
