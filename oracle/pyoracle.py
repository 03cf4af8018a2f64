"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the C FP64 restatement of the reference MNDO path
(oracle/mndo_oracle.c).  Only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
from typing import Dict, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class ScfResult(ctypes.Structure):
    _fields_ = [("e_el", ctypes.c_double), ("e_cc", ctypes.c_double),
                ("e_tot", ctypes.c_double), ("steps", ctypes.c_int),
                ("converged", ctypes.c_int), ("status", ctypes.c_int)]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "mndo_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_n_ao.argtypes = [ctypes.c_int, _ip]
        L.or_pair_block_count.argtypes = [ctypes.c_int, _ip]
        L.or_pair_block_count.restype = ctypes.c_long
        L.or_pair_blocks.argtypes = [ctypes.c_int, _ip, _dp, _dp]
        L.or_get_value.argtypes = [ctypes.c_int, _ip, _dp] + [ctypes.c_int] * 4
        L.or_get_value.restype = ctypes.c_double
        L.or_overlap.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, ctypes.c_int]
        L.or_overlap.restype = ctypes.c_double
        L.or_hcore.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp]
        L.or_guess_density.argtypes = [ctypes.c_int, _ip, _dp]
        L.or_density.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.or_n_electrons.argtypes = [ctypes.c_int, _ip]
        L.or_fock_ref.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        L.or_fock_nddo.argtypes = [ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        L.or_energy.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.or_energy.restype = ctypes.c_double
        L.or_core_core.argtypes = [ctypes.c_int, _ip, _dp, _dp]
        L.or_core_core.restype = ctypes.c_double
        L.or_eigh.argtypes = [ctypes.c_int, _dp, _dp]
        L.or_scf.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                             ctypes.POINTER(ScfResult), _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Molecule:
    """Atoms as arrays: Z (int32) and xyz in Angstrom (n, 3)."""

    def __init__(self, atoms):
        self.Z = np.ascontiguousarray([a[0] for a in atoms], dtype=np.int32)
        self.xyz = np.ascontiguousarray([a[1:4] for a in atoms], dtype=np.float64).reshape(-1, 3)
        self.n_atoms = len(atoms)
        self.n_ao = lib().or_n_ao(self.n_atoms, _i(self.Z))
        self.n_occ = lib().or_n_electrons(self.n_atoms, _i(self.Z)) // 2

    @property
    def T(self) -> int:
        return self.n_ao * (self.n_ao + 1) // 2

    def pair_blocks(self) -> np.ndarray:
        n = lib().or_pair_block_count(self.n_atoms, _i(self.Z))
        out = np.zeros(n)
        lib().or_pair_blocks(self.n_atoms, _i(self.Z), _d(self.xyz), _d(out))
        return out

    def hcore(self, blocks=None) -> np.ndarray:
        blocks = self.pair_blocks() if blocks is None else blocks
        h = np.zeros(self.T)
        lib().or_hcore(self.n_atoms, _i(self.Z), _d(self.xyz), _d(blocks), _d(h))
        return h

    def overlap(self, a: int, b: int) -> float:
        return lib().or_overlap(self.n_atoms, _i(self.Z), _d(self.xyz), a, b)

    def get_value(self, blocks, a, b, c, d) -> float:
        return lib().or_get_value(self.n_atoms, _i(self.Z), _d(blocks), a, b, c, d)

    def guess_density(self) -> np.ndarray:
        p = np.zeros(self.T)
        lib().or_guess_density(self.n_atoms, _i(self.Z), _d(p))
        return p

    def fock(self, blocks, h, p, mode: str = "nddo") -> np.ndarray:
        f = np.zeros(self.T)
        fn = lib().or_fock_ref if mode == "ref" else lib().or_fock_nddo
        fn(self.n_atoms, _i(self.Z), _d(blocks), _d(h), _d(np.ascontiguousarray(p)), _d(f))
        return f

    def core_core(self, blocks=None) -> float:
        blocks = self.pair_blocks() if blocks is None else blocks
        return lib().or_core_core(self.n_atoms, _i(self.Z), _d(self.xyz), _d(blocks))

    def scf(self, threshold: float = 5e-3, max_steps: int = 100, mode: str = "nddo"):
        res = ScfResult()
        en = np.zeros(max_steps + 1)
        p = np.zeros(self.T)
        f = np.zeros(self.T)
        w = np.zeros(self.n_ao)
        lib().or_scf(self.n_atoms, _i(self.Z), _d(self.xyz), threshold, max_steps,
                     0 if mode == "ref" else 1, ctypes.byref(res), _d(en), _d(p), _d(f), _d(w))
        return {"e_el": res.e_el, "e_cc": res.e_cc, "e_tot": res.e_tot, "steps": res.steps,
                "converged": bool(res.converged), "status": res.status,
                "energies": en[: res.steps + 1].copy(), "p": p, "f": f, "eigenvalues": w}


def energy(n, h, f, p) -> float:
    return lib().or_energy(n, _d(np.ascontiguousarray(h)), _d(np.ascontiguousarray(f)),
                           _d(np.ascontiguousarray(p)))


def density(n, n_occ, C_colmajor) -> np.ndarray:
    p = np.zeros(n * (n + 1) // 2)
    c = np.ascontiguousarray(C_colmajor, dtype=np.float64)
    lib().or_density(n, n_occ, _d(c), _d(p))
    return p


def eigh(a_full: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Returns (w ascending, V with eigenvectors as columns)."""
    n = a_full.shape[0]
    buf = np.asfortranarray(a_full, dtype=np.float64).copy(order="F")
    w = np.zeros(n)
    st = lib().or_eigh(n, buf.ctypes.data_as(_dp), _d(w))
    if st != 0:
        raise RuntimeError("or_eigh did not converge (%d)" % st)
    return w, buf


def packed_to_full(p: np.ndarray, n: int) -> np.ndarray:
    out = np.zeros((n, n))
    il = np.tril_indices(n)
    # packed order is row-major over the lower triangle: r(r+1)/2 + c
    rows, cols = il
    order = np.lexsort((cols, rows))
    out[rows[order], cols[order]] = p
    out[cols[order], rows[order]] = p
    return out


def read_dump(path: str) -> Dict[str, np.ndarray]:
    """Reads the record stream written by oracle/ref_dump.cc."""
    out = {}
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        (ln,) = struct.unpack_from("<i", data, pos)
        pos += 4
        name = data[pos:pos + ln].decode()
        pos += ln
        (cnt,) = struct.unpack_from("<q", data, pos)
        pos += 8
        out[name] = np.frombuffer(data, dtype="<f8", count=cnt, offset=pos).copy()
        pos += 8 * cnt
    return out
