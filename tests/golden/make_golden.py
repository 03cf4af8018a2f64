"""Writes the golden fixtures in tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_dump (the reference CPU path built from its own sources
by `make -C oracle ref`) on the shipped geometries of the reference and on the
synthetic workloads, and stores what it printed as compressed .npz files:

  small molecules (N <= 64): every intermediate in full (pair blocks, Hcore,
      overlaps, guess/first Fock, final P/F, eigenvalues, energy trace);
  larger molecules: scalars, energy trace, eigenvalues, the Hcore diagonal and
      a fixed sample of 2048 entries each of the pair blocks, Hcore and the
      first Fock matrix (index + value).

Usage: python tests/golden/make_golden.py [--only NAME ...] [--jobs 8]
Needs /root/reference (this container only); the fixtures travel, the
reference does not.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from semigpu_amd import geometry  # noqa: E402
import pyoracle  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_GEOM = "/root/reference/src/Geometries4Test"
DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
SMALL_N = 64
SAMPLE = 2048


def cases():
    out = {}
    for f in sorted(os.listdir(REF_GEOM)):
        if not f.endswith(".xyz") or f.startswith("filetest"):
            continue
        out["ref_" + f[:-4]] = ("file", os.path.join(REF_GEOM, f))
    out["cfg1_pyridine"] = ("atoms", geometry.pyridine())
    out["cfg2_c40h82"] = ("atoms", geometry.alkane(40))
    for n in (1, 2, 3, 5, 10):
        out["alkane_%02d" % n] = ("atoms", geometry.alkane(n))
    rng = geometry.SplitMix64(99)
    for n in (2, 4, 8, 16, 32):
        out["water_%02d" % n] = ("atoms", geometry.water_cluster(n, rng))
    batch = geometry.water_batch(4096)
    picked = 0
    for i, mol in enumerate(batch):
        if len(mol) // 3 <= 12 and picked < 4:
            out["batch_%04d" % i] = ("atoms", mol)
            picked += 1
    return out


def run_case(name, kind, payload, threshold=1e-6):
    env = dict(os.environ, MKL_INTERFACE_LAYER="LP64", MKL_THREADING_LAYER="SEQUENTIAL",
               OMP_NUM_THREADS="1")
    with tempfile.TemporaryDirectory() as td:
        if kind == "file":
            xyz = payload
        else:
            xyz = os.path.join(td, name + ".xyz")
            with open(xyz, "w") as fh:
                fh.write(geometry.to_xyz(payload, name))
        out = os.path.join(td, name + ".bin")
        subprocess.run([DUMP, xyz, out, repr(threshold)], check=True, env=env,
                       stdout=subprocess.DEVNULL)
        d = pyoracle.read_dump(out)
    n = int(d["n_ao"][0])
    rec = {"Z": d["Z"].astype(np.int32), "xyz_ang": d["xyz_ang"].reshape(-1, 3),
           "n_ao": n, "threshold": threshold}
    for k in ("e_cc", "e_el", "e_tot", "steps", "converged", "shipped_e_el", "shipped_converged",
              "t_integrals_s", "t_hcore_s", "t_scf_s", "t_fock_per_step_s"):
        rec[k] = float(d[k][0])
    rec["energies"] = d["energies"]
    rec["eigenvalues"] = d["eigenvalues"]
    if n <= SMALL_N:
        for k in ("pair_blocks", "hcore", "overlap_offdiag", "guess_f", "iter1_p", "iter1_f",
                  "final_p", "final_f"):
            rec[k] = d[k]
        rec["full"] = 1
    else:
        rs = np.random.default_rng(12345)
        for k in ("pair_blocks", "hcore", "iter1_f"):
            idx = np.sort(rs.choice(d[k].size, size=min(SAMPLE, d[k].size), replace=False))
            rec[k + "_idx"] = idx.astype(np.int64)
            rec[k + "_val"] = d[k][idx]
        rec["hcore_diag"] = d["hcore"][[i * (i + 1) // 2 + i for i in range(n)]]
        rec["full"] = 0
    np.savez_compressed(os.path.join(GOLDEN, name + ".npz"), **rec)
    return name, n, rec["steps"], rec["e_tot"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    if not os.path.exists(DUMP):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    cs = cases()
    if args.only:
        cs = {k: v for k, v in cs.items() if k in args.only}
    with ThreadPoolExecutor(args.jobs) as ex:
        futs = [ex.submit(run_case, k, *v) for k, v in cs.items()]
        for f in futs:
            print(*f.result(), flush=True)
    manifest = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    with open(os.path.join(GOLDEN, "manifest.json"), "w") as fh:
        json.dump({"cases": manifest, "threshold_eV": 1e-6,
                   "source": "oracle/_ref/ref_dump (reference CPU path, g++ -O2, MKL dsyev)"},
                  fh, indent=1)


if __name__ == "__main__":
    main()
