"""CPU checks of the C ABI boundary and of the host-side logic (no GPU calls)."""
import ctypes
import math
import os
import re
import subprocess

import pytest

from conftest import ROOT
from semigpu_amd import api, geometry

HEADER = os.path.join(ROOT, "include", "semigpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(api.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libsemigpu.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", api.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(api.LIB_PATH)
    assert lib.sg_version() >= 100


def test_engine_fails_loudly_without_device(monkeypatch):
    if not os.path.exists(api.LIB_PATH):
        pytest.skip("libsemigpu.so not built")
    monkeypatch.setenv("HIP_VISIBLE_DEVICES", "")
    import torch  # noqa: F401  (only to see whether a device is present here)
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(api.SemiGpuError):
        api.Engine()


def test_missing_library_message(tmp_path):
    saved = api._lib
    api._lib = None
    try:
        with pytest.raises(api.SemiGpuError, match="not built"):
            api.load_library(str(tmp_path / "nope.so"))
    finally:
        api._lib = saved


def _d(a, b):
    return math.dist(a[1:], b[1:])


def test_pyridine_geometry():
    p = geometry.pyridine()
    assert len(p) == 11 and geometry.n_basis(p) == 29 and geometry.n_electrons(p) == 30
    assert abs(_d(p[0], p[1]) - 1.34) < 1e-12 and abs(_d(p[0], p[5]) - 1.34) < 1e-12
    assert abs(_d(p[1], p[2]) - 1.39) < 1e-12 and abs(_d(p[1], p[6]) - 1.08) < 1e-12


def test_alkane_geometry():
    a = geometry.alkane(40)
    assert len(a) == 122 and geometry.n_basis(a) == 242 and geometry.n_electrons(a) == 242
    cs = [x for x in a if x[0] == 6]
    hs = [x for x in a if x[0] == 1]
    for i in range(len(cs) - 1):
        assert abs(_d(cs[i], cs[i + 1]) - 1.54) < 1e-12
    for h in hs:
        assert abs(min(_d(h, c) for c in cs) - 1.09) < 1e-12
    for i in range(1, len(cs) - 1):
        u = [cs[i - 1][k] - cs[i][k] for k in (1, 2, 3)]
        v = [cs[i + 1][k] - cs[i][k] for k in (1, 2, 3)]
        ang = math.degrees(math.acos(sum(x * y for x, y in zip(u, v)) / (1.54 * 1.54)))
        assert abs(ang - 109.47) < 1e-9
    m = geometry.alkane(1)
    assert len(m) == 5 and all(abs(_d(m[0], h) - 1.09) < 1e-12 for h in m[1:])


def test_water_batch_deterministic_and_shaped():
    b1 = geometry.water_batch(64, seed=3)
    b2 = geometry.water_batch(64, seed=3)
    assert b1 == b2
    assert geometry.water_batch(64, seed=4) != b1
    for mol in b1:
        n = len(mol) // 3
        assert 2 <= n <= 64
        for w in range(n):
            o, h1, h2 = mol[3 * w: 3 * w + 3]
            assert abs(_d(o, h1) - 0.96) < 1e-12 and abs(_d(o, h2) - 0.96) < 1e-12
            u = [h1[k] - o[k] for k in (1, 2, 3)]
            v = [h2[k] - o[k] for k in (1, 2, 3)]
            ang = math.degrees(math.acos(sum(x * y for x, y in zip(u, v)) / 0.96 ** 2))
            assert abs(ang - 104.5) < 1e-9


def test_rng_transforms():
    r = geometry.SplitMix64(1)
    xs = [r.uniform() for _ in range(20000)]
    assert 0.0 <= min(xs) and max(xs) < 1.0 and abs(sum(xs) / len(xs) - 0.5) < 0.01
    ks = [r.randint(2, 64) for _ in range(20000)]
    assert min(ks) == 2 and max(ks) == 64
    rot = r.rotation()
    for i in range(3):
        for j in range(3):
            dot = sum(rot[i][k] * rot[j][k] for k in range(3))
            assert abs(dot - (1.0 if i == j else 0.0)) < 1e-12


def test_serving_stream_shape():
    s = geometry.serving_stream(200.0, 1.0, seed=2)
    assert 150 < len(s) < 260
    assert all(t2 >= t1 for (t1, _, _), (t2, _, _) in zip(s, s[1:]))
