"""Continuous-batching server (sg_server) through the C ABI.

A molecule's result must not depend on what else is in flight: every result
from the server is compared bit for bit with the same molecule run alone as a
batch of one, and the reference fixtures are met to the SCF bars
(|E - E_ref| <= 1e-6 eV at threshold 1e-6 eV, same step count).
"""
import sys

import numpy as np
import pytest

from conftest import atoms_of, golden_names, load_golden
from semigpu_amd import api, geometry

pytestmark = pytest.mark.gpu

KEYS = ("e_el", "e_cc", "e_tot", "steps", "converged", "status")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def stream():
    """Mixed sizes: small water clusters, a few large ones, alkanes, pyridine."""
    mols = geometry.water_batch(24, seed=314, n_min=2, n_max=24)
    mols[5:5] = [geometry.water_cluster(60, geometry.SplitMix64(9))]
    mols[11:11] = [geometry.alkane(9)]
    mols[17:17] = [geometry.water_cluster(48, geometry.SplitMix64(10)), geometry.pyridine()]
    return mols


def test_server_equals_alone_bitwise(engine):
    mols = stream()
    # small pools: the large clusters wait for space, ranges and slots are reused
    srv = engine.server(hbm_bytes=48e6, max_slots=6, max_atoms=600, threshold=1e-6)
    for i, m in enumerate(mols):
        srv.submit(1000 + i, m)
    out = srv.drain()
    srv.close()
    assert sorted(r["id"] for r in out) == [1000 + i for i in range(len(mols))]
    by_id = {r["id"]: r for r in out}
    for i, m in enumerate(mols):
        r = by_id[1000 + i]
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]
        alone = engine.batch([m]).scf(threshold=1e-6)[0]
        say("server", i, len(m), r["steps"], r["e_tot"], alone["e_tot"])
        assert {k: r[k] for k in KEYS} == alone, i


def test_server_mixed_policy_equals_alone(engine):
    mols = geometry.water_batch(16, seed=2718, n_min=2, n_max=32)
    srv = engine.server(hbm_bytes=64e6, max_slots=5, max_atoms=400, threshold=1e-6,
                        policy="mixed")
    for i, m in enumerate(mols):
        srv.submit(i, m)
    out = {r["id"]: r for r in srv.drain()}
    srv.close()
    for i, m in enumerate(mols):
        alone = engine.batch([m]).scf(threshold=1e-6, policy="mixed")[0]
        assert {k: out[i][k] for k in KEYS} == alone, i


def test_server_mixed_policy_large_molecules_equal_alone(engine):
    """Mixed policy with molecules above N = 384 (wide reduction panels, FP32
    steps in the combined launch) in one class with small ones."""
    mols = [geometry.water_cluster(80, geometry.SplitMix64(80)),      # N = 480
            geometry.water_cluster(110, geometry.SplitMix64(110))]    # N = 660
    mols += geometry.water_batch(6, seed=31, n_min=2, n_max=20)
    srv = engine.server(hbm_bytes=512e6, max_slots=8, max_atoms=1024, threshold=1e-6,
                        policy="mixed")
    for i, m in enumerate(mols):
        srv.submit(i, m)
    out = {r["id"]: r for r in srv.drain()}
    srv.close()
    for i, m in enumerate(mols):
        alone = engine.batch([m]).scf(threshold=1e-6, policy="mixed")[0]
        say("mixed server", i, geometry.n_basis(m), out[i]["steps"], alone["steps"])
        assert {k: out[i][k] for k in KEYS} == alone, i


def test_server_fixtures(engine):
    names = golden_names()
    srv = engine.server(hbm_bytes=256e6, max_slots=16, max_atoms=4096, threshold=1e-6)
    for i, n in enumerate(names):
        srv.submit(i, atoms_of(load_golden(n)))
    out = {r["id"]: r for r in srv.drain()}
    srv.close()
    for i, n in enumerate(names):
        g = load_golden(n)
        r = out[i]
        assert r["steps"] == int(g["steps"]), n
        assert r["converged"] == bool(g["converged"]), n
        assert abs(r["e_tot"] - float(g["e_tot"])) <= 1e-6, n


def test_server_rejects(engine):
    srv = engine.server(hbm_bytes=16e6, max_slots=4, max_atoms=256)
    with pytest.raises(api.SemiGpuError, match="Z=9"):
        srv.submit(1, [(9, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 0.9)])
    with pytest.raises(api.SemiGpuError, match="capacity"):
        srv.submit(2, geometry.water_cluster(100, geometry.SplitMix64(1)))  # 300 atoms > max_atoms
    assert srv.step() == (0, 0)
    assert srv.poll() == []
    srv.close()


def test_server_step_cap(engine):
    srv = engine.server(hbm_bytes=16e6, max_slots=4, max_atoms=256, threshold=1e-12,
                        max_steps=3)
    srv.submit(7, geometry.pyridine())
    out = srv.drain()
    srv.close()
    assert len(out) == 1 and out[0]["steps"] == 3 and not out[0]["converged"]
    alone = engine.batch([geometry.pyridine()]).scf(threshold=1e-12, max_steps=3)[0]
    assert {k: out[0][k] for k in KEYS} == alone


def test_size_classes_equal_alone_and_decouple(engine):
    """Size classes: every result is the molecule's result alone, and small
    molecules queued behind a large one finish before it (their class
    iterates at its own pace on its own stream)."""
    big = geometry.water_cluster(100, geometry.SplitMix64(21))       # N = 600
    small = geometry.water_batch(12, seed=5, n_min=2, n_max=12)      # N <= 72
    srv = engine.server(hbm_bytes=512e6, max_slots=32, max_atoms=2048, threshold=1e-6,
                        classes=(200,))
    srv.submit(0, big)
    for i, m in enumerate(small):
        srv.submit(1 + i, m)
    out = {r["id"]: r for r in srv.drain()}
    srv.close()
    assert sorted(out) == list(range(len(small) + 1))
    say("big done", out[0]["t_done"], "small done", max(out[1 + i]["t_done"]
                                                        for i in range(len(small))))
    assert max(out[1 + i]["t_done"] for i in range(len(small))) < out[0]["t_done"]
    for i, m in enumerate([big] + small):
        alone = engine.batch([m]).scf(threshold=1e-6)[0]
        assert {k: out[i][k] for k in KEYS} == alone, i


def test_size_class_bounds_rejected(engine):
    with pytest.raises(api.SemiGpuError, match="bounds"):
        engine.server(hbm_bytes=64e6, max_slots=8, max_atoms=512, classes=(300, 200))


def test_server_edge_molecules(engine):
    """Single atoms, H2 and OH through the server (one class, then alone)."""
    mols = [[(1, 0.0, 0.0, 0.0)], [(8, 0.0, 0.0, 0.0)], [(1, 0.0, 0.0, 0.0), (1, 0.0, 0.0, 0.74)],
            [(1, 0.0, 0.0, 0.0), (8, 0.0, 0.0, 0.97)]]
    srv = engine.server(hbm_bytes=16e6, max_slots=4, max_atoms=64, threshold=1e-6)
    srv.submit(0, mols[0])
    out = srv.drain()                      # a lone H atom in flight (N = 1)
    for i, m in enumerate(mols[1:], 1):
        srv.submit(i, m)
    out += srv.drain()
    srv.close()
    by_id = {r["id"]: r for r in out}
    for i, m in enumerate(mols):
        assert {k: by_id[i][k] for k in KEYS} == engine.batch([m]).scf(threshold=1e-6)[0], i


def test_server_diis_equals_alone(engine):
    """DIIS in the server: every result equals the molecule alone as a batch of
    one with the same DIIS depth, bit for bit (FP64 and mixed policy), in no
    more steps than the plain SCF and at its energy; the depth is only set
    while the server is idle."""
    mols = stream()
    for policy in ("fp64", "mixed"):
        srv = engine.server(hbm_bytes=48e6, max_slots=6, max_atoms=600, threshold=1e-6,
                            policy=policy)
        srv.set_diis(6)
        for i, m in enumerate(mols):
            srv.submit(i, m)
        with pytest.raises(api.SemiGpuError, match="in flight"):
            srv.set_diis(4)
        out = {r["id"]: r for r in srv.drain()}
        srv.set_diis(0)
        srv.close()
        assert sorted(out) == list(range(len(mols)))
        for i, m in enumerate(mols):
            b = engine.batch([m])
            plain = b.scf(threshold=1e-6, policy=policy)[0]
            b.set_diis(6)
            alone = b.scf(threshold=1e-6, policy=policy)[0]
            b.close()
            say("server diis", policy, i, out[i]["steps"], "plain", plain["steps"],
                out[i]["e_tot"] - plain["e_tot"])
            assert {k: out[i][k] for k in KEYS} == alone, (policy, i)
            assert alone["converged"] and alone["steps"] <= plain["steps"]
            assert abs(alone["e_tot"] - plain["e_tot"]) <= 1e-5


def serving_slice():
    """A seeded slice of the serving stream (200/s, seed 7): its first 12
    requests (alkanes up to C46H94, water clusters up to (H2O)161, N = 966),
    the largest water cluster of the first 5 s ((H2O)199, N = 1194), the
    longest alkane (C60H122) and the alkanes that have reference fixtures."""
    reqs = geometry.serving_stream(200.0, 5.0, 7)
    picks = list(range(12)) + [108, 47, 22, 41, 72, 82, 98]
    return [(i, reqs[i][1], reqs[i][2]) for i in picks]


FIXTURE_OF = {"C2H6": "alkane_02", "C3H8": "alkane_03", "C5H12": "alkane_05",
              "C10H22": "alkane_10", "C40H82": "cfg2_c40h82"}


@pytest.mark.parametrize("policy", ["fp64", "mixed"])
def test_serving_stream_slice_three_classes(engine, policy):
    """The serving stream's molecules through the server with the bench's
    three size classes (N < 300, < 700, >= 700): two water clusters above
    N = 700 in the large class, every result equal to the molecule alone bit
    for bit, and the alkanes with reference fixtures on the fixtures (FP64:
    same steps, 1e-6 eV; mixed: 1e-4 eV)."""
    picks = serving_slice()
    assert sum(1 for _, _, m in picks if geometry.n_basis(m) >= 700) >= 2
    srv = engine.server(hbm_bytes=4e9, max_slots=64, max_atoms=1 << 14, threshold=1e-6,
                        policy=policy, classes=(300, 700))
    for i, label, m in picks:
        say("submit", policy, i, label, geometry.n_basis(m))
        srv.submit(i, m)
    out = {r["id"]: r for r in srv.drain()}
    srv.close()
    assert sorted(out) == sorted(i for i, _, _ in picks)
    for i, label, m in picks:
        r = out[i]
        alone = engine.batch([m]).scf(threshold=1e-6, policy=policy)[0]
        say("stream", policy, i, label, r["n_ao"], r["steps"], r["e_tot"], alone["e_tot"])
        assert r["converged"] and r["n_ao"] == geometry.n_basis(m)
        assert {k: r[k] for k in KEYS} == alone, (policy, i, label)
        if label in FIXTURE_OF:
            g = load_golden(FIXTURE_OF[label])
            if policy == "fp64":
                assert r["steps"] == int(g["steps"]), label
                assert abs(r["e_tot"] - float(g["e_tot"])) <= 1e-6, label
            else:
                assert abs(r["e_tot"] - float(g["e_tot"])) <= 1e-4, label
