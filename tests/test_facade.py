"""CPU checks of the C++ classes with the reference's names (semigpu_amd/cpp).

* libsemigpu_cpp.so exports the reference's class members;
* the reference's own main (src/App/semiGPU.cc) compiles and links unchanged
  against these headers (needs the reference sources, so it runs here only);
* the XYZ reader reads every shipped geometry exactly as the reference reader
  did when the fixtures were written (Z and Angstrom coordinates bit for bit),
  and the reference's reader test files as the compiled reference test does.
No device call is made: ReadXYZFile and Atom are host-side bookkeeping.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_names, load_golden

CPP = os.path.join(ROOT, "semigpu_amd", "cpp")
LIB = os.path.join(ROOT, "semigpu_amd", "libsemigpu_cpp.so")
XYZ_PROBE = os.path.join(ROOT, "tests", "cpp", "xyz_probe")
REF = "/root/reference/src"
REF_GEOM = os.path.join(REF, "Geometries4Test")
REF_READER_TEST = os.path.join(ROOT, "oracle", "_ref", "test_readxyzfile")

MEMBERS = [
    "TwoCenterIntegral::ComputeAllTwoCenterIntegral", "TwoCenterIntegral::GetValueFromArray",
    "TwoCenterIntegral::ComputeTwoCenterIntegral", "TwoCenterIntegral::Alloc4AllTwoCenterIntegral",
    "TwoCenterIntegral::Dealloc4AllTwoCenterIntegral", "TwoCenterIntegral::To_device",
    "Hcore::ComputeMatrix", "Hcore::ComputeElementMatrix", "BaseMatrix::Alloc4Matrix",
    "BaseMatrix::Dealloc4Matrix", "DensityMatrix::GuessDensityMatrixSwitch",
    "DensityMatrix::ComputeMatrix", "FockMatrix::ComputeMatrix", "SCFCalculation::ComputeSCF",
    "SCFCalculation::AllocSCFData", "SCFCalculation::DeallocSCFData",
    "ElectronicEnergy::ComputeEnergy", "CoreCoreRepulsion::ComputeRepulsion",
    "ReadXYZFile::GetValuesFromFile", "ListAtomicOrbitals::SetOrbitals", "Atom::PrintGeometry",
    "ScreenUtils::PrintScrStarLine",
]


def need(path):
    if not os.path.exists(path):
        pytest.skip("%s not built (run __graft_entry__.build())" % path)


def probe(path):
    need(XYZ_PROBE)
    out = subprocess.run([XYZ_PROBE, path], capture_output=True, text=True, check=True).stdout
    lines = out.strip().splitlines()
    k = next(i for i, l in enumerate(lines) if l.split()[0] in ("ok", "fail"))
    ok, n = lines[k].split()
    atoms = [l.split() for l in lines[k + 1:]]
    return ok == "ok", int(n), atoms


def test_library_exports_reference_members():
    need(LIB)
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    missing = [m for m in MEMBERS if m + "(" not in out]
    assert not missing, missing


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources absent")
def test_reference_main_builds_against_classes(tmp_path):
    need(LIB)
    exe = tmp_path / "semiGPU_ref_main"
    subprocess.run(["g++", "-std=c++17", "-I", CPP, "-o", str(exe),
                    os.path.join(REF, "App", "semiGPU.cc"), "-L", os.path.dirname(LIB),
                    "-lsemigpu_cpp", "-lsemigpu"], check=True, capture_output=True)
    assert exe.exists()


def test_reader_formats(tmp_path):
    f = tmp_path / "a.xyz"
    f.write_text("3\ncomment\nC 0.0 0.0 0.0\n1 1.09 0 0\nO 0 1.2 -0.5\n")
    ok, n, atoms = probe(str(f))
    assert ok and n == 3
    assert [a[0] for a in atoms] == ["6", "1", "8"]
    assert [a[5] for a in atoms] == ["4", "1", "6"]  # valence = core charge
    assert [float(x) for x in atoms[2][2:5]] == [0.0, 1.2, -0.5]
    f.write_text("2\n\nXx 0 0 0\nH 0 0 1\n")
    ok, n, atoms = probe(str(f))
    assert not ok and atoms[0][6] == "0"
    f.write_text("0\nempty\n")
    assert not probe(str(f))[0]
    assert not probe(str(tmp_path / "missing.xyz"))[0]


def test_reader_concatenated_records(tmp_path):
    need(XYZ_PROBE)
    f = tmp_path / "many.xyz"
    f.write_text("1\nH atom\nH 0 0 0\n\n2\nH2\nH 0 0 0\nH 0 0 0.74\n3\nwater\n"
                 "O 0 0 0\nH 0.96 0 0\nH -0.24 0.93 0\n")
    out = subprocess.run([XYZ_PROBE, "-r", str(f)], capture_output=True, text=True,
                         check=True).stdout.split()
    assert out == ["ok", "3", "1", "2", "3"]


@pytest.mark.skipif(not os.path.isdir(REF_GEOM), reason="reference geometries absent")
@pytest.mark.parametrize("name", [n for n in golden_names() if n.startswith("ref_")])
def test_reader_matches_reference_reader(name):
    """The fixtures' Z and coordinates were read by the reference's ReadXYZFile."""
    g = load_golden(name)
    ok, n, atoms = probe(os.path.join(REF_GEOM, name[4:] + ".xyz"))
    assert ok and n == len(g["Z"])
    assert [int(a[0]) for a in atoms] == [int(z) for z in g["Z"]]
    xyz = np.array([[float(x) for x in a[2:5]] for a in atoms])
    assert np.array_equal(xyz, g["xyz_ang"])


@pytest.mark.skipif(not os.path.exists(REF_READER_TEST), reason="reference reader test not built")
@pytest.mark.parametrize("f", ["filetest_4_readxyz_class_1.xyz", "filetest_4_readxyz_class_2.xyz"])
def test_reader_test_files_like_reference(f):
    path = os.path.join(REF_GEOM, f)
    out = subprocess.run([REF_READER_TEST, path], capture_output=True, text=True).stdout
    lines = out.splitlines()
    k = next(i for i, l in enumerate(lines) if l.startswith("All data is fine ?"))
    ref_ok = lines[k].split()[-1] == "yes"
    ref_atoms = [l.split() for l in lines[k + 1:] if l.strip()]
    ok, n, atoms = probe(path)
    assert ok == ref_ok
    assert [int(a[0]) for a in atoms] == [int(a[0]) for a in ref_atoms]
    mine = np.array([[float(x) for x in a[2:5]] for a in atoms])
    theirs = np.array([[float(x) for x in a[1:4]] for a in ref_atoms])
    assert np.allclose(mine, theirs, rtol=1e-5, atol=1e-6)  # reference prints 6 digits
