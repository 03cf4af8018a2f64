"""Python binding of the C ABI (include/semigpu.h) over the in-tree libsemigpu.so.

There is no CPU fallback: importing works anywhere, but creating an
:class:`Engine` needs the built library and a HIP device, and raises
otherwise.

    eng = Engine()                         # GPU 0, one stream
    b = eng.batch([geometry.pyridine()])   # list of molecules (Z, x, y, z) in Angstrom
    res = b.scf(threshold=1e-6)            # SCFCalculation::ComputeSCF for every molecule
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from typing import List, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsemigpu.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "semigpu.h")

SG_HCORE, SG_FOCK, SG_DENSITY = 0, 1, 2
STAGES = ("two_center", "hcore", "guess", "fock", "tridiag", "bisect", "invit", "backtransform",
          "density", "energy", "control", "tridiag_fp32", "tridiag_head")
EIGH_STAGES = ("tridiag_head", "tridiag", "tridiag_fp32", "bisect", "invit", "backtransform")
TUNE_BUCKETS = 8       # SG_TUNE_BUCKETS
TUNE_CANDIDATES = 3    # SG_TUNE_CANDIDATES

# Every function declared in include/semigpu.h (checked by tests/test_abi.py).
EXPORTS = (
    "sg_last_error", "sg_version", "sg_ctx_create", "sg_ctx_destroy", "sg_batch_create",
    "sg_batch_destroy", "sg_batch_molecule_info", "sg_two_center", "sg_hcore", "sg_core_core",
    "sg_guess_density", "sg_fock", "sg_eigh", "sg_density", "sg_energy", "sg_scf_run",
    "sg_scf_run_policy", "sg_scf_fp32_steps", "sg_scf_set_damping", "sg_scf_set_diis", "sg_gradient_fd",
    "sg_pair_block_count", "sg_get_pair_blocks", "sg_get_packed", "sg_set_packed",
    "sg_get_eigen", "sg_set_eigen", "sg_batch_create_bohr", "sg_set_pair_blocks",
    "sg_get_status", "sg_calibration_copy", "sg_timing_enable", "sg_timing_read", "sg_timing_reset",
    "sg_server_create", "sg_server_destroy", "sg_server_submit", "sg_server_step",
    "sg_server_poll", "sg_server_clock", "sg_server_set_diis", "sg_ctx_set_large_basis", "sg_server_create_classes",
    "sg_heat_of_formation", "sg_ctx_set_graphs", "sg_ctx_set_overlap", "sg_eigh_noalloc", "sg_gradient",
    "sg_scf_record_energies", "sg_scf_energy_history", "sg_ctx_autotune",
    "sg_ctx_set_tri_variant", "sg_ctx_tri_shapes",
)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class ScfResult(ctypes.Structure):
    _fields_ = [("e_el", ctypes.c_double), ("e_cc", ctypes.c_double), ("e_tot", ctypes.c_double),
                ("steps", ctypes.c_int), ("converged", ctypes.c_int), ("status", ctypes.c_int)]


class Served(ctypes.Structure):
    _fields_ = [("request_id", ctypes.c_longlong), ("e_el", ctypes.c_double),
                ("e_cc", ctypes.c_double), ("e_tot", ctypes.c_double), ("steps", ctypes.c_int),
                ("converged", ctypes.c_int), ("status", ctypes.c_int), ("n_ao", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]


class SemiGpuError(RuntimeError):
    pass


# precision policies of sg_scf_run_policy / sg_server_create (include/semigpu.h)
POLICIES = {"fp64": 0, "mixed": 1, "mixed_bf16": 2, "mixed_eig32": 3}


_lib = None
_engines = weakref.WeakSet()


@atexit.register
def _close_engines():
    # Engines (and their batches and servers) still open at interpreter exit
    # are destroyed before module teardown, while the HIP runtime is intact.
    for e in list(_engines):
        e.close()


def load_library(path: str = LIB_PATH):
    """Loads libsemigpu.so and declares its signatures; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SemiGpuError("%s is not built: run `python -m semigpu_amd.build` "
                           "(or __graft_entry__.build())" % path)
    L = ctypes.CDLL(path)
    L.sg_last_error.restype = ctypes.c_char_p
    L.sg_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.sg_ctx_destroy.argtypes = [_vp]
    L.sg_ctx_set_large_basis.argtypes = [_vp, ctypes.c_int]
    L.sg_ctx_set_graphs.argtypes = [_vp, ctypes.c_int]
    L.sg_ctx_set_overlap.argtypes = [_vp, ctypes.c_int]
    L.sg_gradient.argtypes = [_vp, _dp]
    L.sg_scf_record_energies.argtypes = [_vp, ctypes.c_int]
    L.sg_ctx_autotune.argtypes = [_vp, ctypes.c_int, _dp, _ip]
    L.sg_ctx_set_tri_variant.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    L.sg_ctx_tri_shapes.argtypes = [_vp, _ip, _ip, _ip, _ip]
    L.sg_scf_energy_history.argtypes = [_vp, ctypes.c_int, _dp, _ip]
    L.sg_eigh_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]
    L.sg_heat_of_formation.argtypes = [ctypes.c_int, _ip, ctypes.c_double, _dp]
    L.sg_batch_create.argtypes = [_vp, ctypes.c_int, _ip, _ip, _dp, ctypes.POINTER(_vp)]
    L.sg_batch_destroy.argtypes = [_vp]
    L.sg_batch_molecule_info.argtypes = [_vp, ctypes.c_int, _ip, _ip, _ip]
    for f in ("sg_two_center", "sg_hcore", "sg_guess_density", "sg_fock", "sg_eigh", "sg_density"):
        getattr(L, f).argtypes = [_vp]
    L.sg_core_core.argtypes = [_vp, _dp]
    L.sg_energy.argtypes = [_vp, _dp]
    L.sg_scf_run.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ScfResult)]
    L.sg_scf_run_policy.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.POINTER(ScfResult)]
    L.sg_scf_fp32_steps.argtypes = [_vp, _ip]
    L.sg_scf_set_damping.argtypes = [_vp, ctypes.c_double]
    L.sg_scf_set_diis.argtypes = [_vp, ctypes.c_int]
    L.sg_gradient_fd.argtypes = [_vp, ctypes.c_int, _ip, _dp, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, _dp, _dp, _ip]
    L.sg_server_set_diis.argtypes = [_vp, ctypes.c_int]
    L.sg_pair_block_count.argtypes = [_vp, ctypes.c_int]
    L.sg_pair_block_count.restype = ctypes.c_long
    L.sg_get_pair_blocks.argtypes = [_vp, ctypes.c_int, _dp]
    L.sg_get_packed.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp]
    L.sg_set_packed.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp]
    L.sg_get_eigen.argtypes = [_vp, ctypes.c_int, _dp, _dp]
    L.sg_set_eigen.argtypes = [_vp, ctypes.c_int, _dp, _dp]
    L.sg_set_pair_blocks.argtypes = [_vp, ctypes.c_int, _dp]
    L.sg_get_status.argtypes = [_vp, _ip]
    L.sg_batch_create_bohr.argtypes = [_vp, ctypes.c_int, _ip, _ip, _dp, ctypes.POINTER(_vp)]
    L.sg_timing_enable.argtypes = [_vp, ctypes.c_int]
    L.sg_calibration_copy.argtypes = [_vp, ctypes.c_longlong]
    L.sg_timing_read.argtypes = [_vp, _dp, _ip]
    L.sg_timing_reset.argtypes = [_vp]
    L.sg_server_create.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.POINTER(_vp)]
    L.sg_server_create_classes.argtypes = [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int, _ip,
                                           ctypes.POINTER(_vp)]
    L.sg_server_destroy.argtypes = [_vp]
    L.sg_server_submit.argtypes = [_vp, ctypes.c_longlong, ctypes.c_int, _ip, _dp]
    L.sg_server_step.argtypes = [_vp, _ip, _ip]
    L.sg_server_poll.argtypes = [_vp, ctypes.POINTER(Served), ctypes.c_int]
    L.sg_server_clock.argtypes = [_vp]
    L.sg_server_clock.restype = ctypes.c_double
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise SemiGpuError("semigpu error %d: %s" % (rc, _lib.sg_last_error().decode()))


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def heat_of_formation(molecule, e_tot: float) -> float:
    """MNDO heat of formation (kcal/mol) of a molecule [(Z, x, y, z), ...] from
    its SCF total energy in eV (sg_heat_of_formation; host arithmetic)."""
    L = load_library()
    Z = np.asarray([int(a[0]) for a in molecule], dtype=np.int32)
    out = ctypes.c_double()
    _check(L.sg_heat_of_formation(len(Z), _i(Z), float(e_tot), ctypes.byref(out)))
    return out.value


class Engine:
    """One HIP device + one stream (sg_context)."""

    def __init__(self, device: int = 0):
        L = load_library()
        h = _vp()
        _check(L.sg_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.L = L
        self._batches = weakref.WeakSet()
        self._servers = weakref.WeakSet()
        _engines.add(self)

    def close(self):
        """Destroys the context; batches still alive are destroyed first."""
        if self.h:
            for b in list(self._batches):
                b.close()
            for srv in list(self._servers):
                srv.close()
            self.L.sg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def batch(self, molecules: Sequence[Sequence[Sequence[float]]] = None, packed=None) -> "Batch":
        return Batch(self, molecules, packed)

    def server(self, **kw) -> "Server":
        return Server(self, **kw)

    def gradient_fd(self, molecule, step: float = 1e-3, threshold: float = 1e-9,
                    diis: int = 6):
        """Nuclear gradient (eV/A, n_atoms x 3) of the MNDO total energy by
        central differences over one batched SCF of the 6 n_atoms displaced
        geometries (sg_gradient_fd).  Returns (gradient, e_tot, all converged)."""
        Z = np.ascontiguousarray([int(a[0]) for a in molecule], dtype=np.int32)
        xyz = np.ascontiguousarray([a[1:4] for a in molecule], dtype=np.float64).reshape(-1)
        g = np.zeros(3 * len(Z))
        e = ctypes.c_double()
        conv = ctypes.c_int()
        _check(self.L.sg_gradient_fd(self.h, len(Z), _i(Z), _d(xyz), float(step), float(threshold),
                                     int(diis), ctypes.byref(e), _d(g), ctypes.byref(conv)))
        return g.reshape(-1, 3), e.value, bool(conv.value)

    def gradient(self, molecule, threshold: float = 1e-10, diis: int = 6, max_steps: int = 200,
                 polish: int = 15):
        """Analytic nuclear gradient (eV/A, n_atoms x 3) of the MNDO total energy
        of one molecule -> (gradient, E_tot, converged).  The derivatives at
        fixed density (sg_gradient) are exact at self-consistency and off by
        the first order of the SCF residual elsewhere, so after the SCF to
        `threshold` (DIIS of depth `diis`, 0 = off) the density is polished by
        repeating the SCF for `polish` steps beyond its step count with no
        energy test (threshold 0): about 1e-8 eV/A instead of about 1e-6 at a
        1e-12 eV energy test."""
        b = self.batch([molecule])
        if diis:
            b.set_diis(diis)
        r = b.scf(threshold=threshold, max_steps=max_steps)[0]
        if polish:
            b.scf(threshold=0.0, max_steps=r["steps"] + polish)
        g = b.gradient()
        b.close()
        return g, r["e_tot"], bool(r["converged"] and r["status"] == 0)

    def autotune(self, molecules: int = 512) -> dict:
        """Times every launch shape of every size bucket of the on-chip FP64
        reduction (sg_ctx_autotune) and keeps the fastest for later batches
        and servers.  Returns {"ms": buckets x candidates (-1: none),
        "choice": per bucket, "shapes": per bucket [(wavefronts, tiles per
        wavefront), ...]}."""
        nb, nc = TUNE_BUCKETS, TUNE_CANDIDATES
        ms = np.zeros(nb * nc)
        ch = np.zeros(nb, dtype=np.int32)
        _check(self.L.sg_ctx_autotune(self.h, int(molecules), _d(ms), _i(ch)))
        out = self.tri_shapes()
        out["ms"] = ms.reshape(nb, nc).round(4).tolist()
        return out

    def tri_shapes(self) -> dict:
        """The reduction buckets' launch shapes and the current choice."""
        nb, nc = TUNE_BUCKETS, TUNE_CANDIDATES
        var = np.zeros(nb, dtype=np.int32)
        cnt = np.zeros(nb, dtype=np.int32)
        nw = np.zeros(nb * nc, dtype=np.int32)
        nr = np.zeros(nb * nc, dtype=np.int32)
        _check(self.L.sg_ctx_tri_shapes(self.h, _i(var), _i(cnt), _i(nw), _i(nr)))
        shapes = [[(int(nw[b * nc + c]), int(nr[b * nc + c])) for c in range(int(cnt[b]))]
                  for b in range(nb)]
        return {"choice": var.tolist(), "shapes": shapes}

    def set_tri_variant(self, bucket: int, variant: int):
        """Forces one reduction bucket's launch shape (sg_ctx_set_tri_variant)."""
        _check(self.L.sg_ctx_set_tri_variant(self.h, int(bucket), int(variant)))

    def set_large_basis(self, n_min: int):
        """Molecules of later batches with N >= n_min take the multi-workgroup
        eigensolver (default: only those above the one-workgroup limit, 1280)."""
        _check(self.L.sg_ctx_set_large_basis(self.h, int(n_min)))

    def eigh_noalloc(self, n: int, a_ptr: int, eigval_ptr: int, eigvec_ptr: int,
                     stream: int = 0):
        """sg_eigh_noalloc on device pointers (integers): queues the full
        eigendecomposition of the n x n matrix at a_ptr on `stream` (0: the
        context's) without allocating; the caller orders and waits."""
        _check(self.L.sg_eigh_noalloc(self.h, int(n), _vp(a_ptr), _vp(eigval_ptr),
                                      _vp(eigvec_ptr), _vp(stream) if stream else None))

    def graphs(self, on: bool = True):
        """Capture SCF iterations as HIP graphs (default) or launch them directly."""
        _check(self.L.sg_ctx_set_graphs(self.h, 1 if on else 0))

    def overlap(self, on: bool = True):
        """Iterate batches with molecules on both sides of 256 basis functions
        as two groups on two streams (default), or as one group (each kernel
        alone on the device)."""
        _check(self.L.sg_ctx_set_overlap(self.h, 1 if on else 0))

    def timing(self, on: bool = True):
        _check(self.L.sg_timing_enable(self.h, 1 if on else 0))

    def timing_read(self):
        ms = np.zeros(len(STAGES))
        cnt = np.zeros(len(STAGES), dtype=np.int32)
        _check(self.L.sg_timing_read(self.h, _d(ms), _i(cnt)))
        return {s: (float(ms[k]), int(cnt[k])) for k, s in enumerate(STAGES)}

    def timing_reset(self):
        _check(self.L.sg_timing_reset(self.h))

    def calibration_copy(self, nbytes: int):
        """One plain copy kernel reading and writing nbytes (byte-counter scale)."""
        _check(self.L.sg_calibration_copy(self.h, int(nbytes)))


def pack(molecules):
    """Molecules [(Z, x, y, z), ...] (Angstrom) -> (atom offsets, Z, xyz), the
    arrays sg_batch_create takes."""
    offs = [0]
    Z: List[int] = []
    xyz: List[float] = []
    for mol in molecules:
        for a in mol:
            Z.append(int(a[0]))
            xyz.extend(float(c) for c in a[1:4])
        offs.append(len(Z))
    return (np.asarray(offs, dtype=np.int32), np.asarray(Z, dtype=np.int32),
            np.asarray(xyz, dtype=np.float64))


class Batch:
    """A batch of molecules resident on the device (sg_batch)."""

    def __init__(self, engine: Engine, molecules=None, packed=None):
        """molecules [(Z, x, y, z), ...] each, or packed = pack(molecules)."""
        self.engine = engine
        self.L = engine.L
        self.offsets, self.Z, self.xyz = packed if packed is not None else pack(molecules)
        self.n_mol = len(self.offsets) - 1
        h = _vp()
        _check(self.L.sg_batch_create(engine.h, self.n_mol, _i(self.offsets), _i(self.Z),
                                      _d(self.xyz), ctypes.byref(h)))
        self.h = h
        engine._batches.add(self)
        self._info = None

    @property
    def info(self):
        """(atoms, basis functions, occupied orbitals) per molecule."""
        if self._info is None:
            info = []
            for m in range(self.n_mol):
                na, nao, nocc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
                _check(self.L.sg_batch_molecule_info(self.h, m, ctypes.byref(na),
                                                     ctypes.byref(nao), ctypes.byref(nocc)))
                info.append((na.value, nao.value, nocc.value))
            self._info = info
        return self._info

    def close(self):
        if self.h:
            self.L.sg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n_ao(self, m: int) -> int:
        return self.info[m][1]

    # stages -------------------------------------------------------------
    def two_center(self):
        _check(self.L.sg_two_center(self.h))

    def hcore(self):
        _check(self.L.sg_hcore(self.h))

    def core_core(self) -> np.ndarray:
        out = np.zeros(self.n_mol)
        _check(self.L.sg_core_core(self.h, _d(out)))
        return out

    def guess_density(self):
        _check(self.L.sg_guess_density(self.h))

    def fock(self):
        _check(self.L.sg_fock(self.h))

    def eigh(self):
        _check(self.L.sg_eigh(self.h))

    def density(self):
        _check(self.L.sg_density(self.h))

    def energy(self) -> np.ndarray:
        out = np.zeros(self.n_mol)
        _check(self.L.sg_energy(self.h, _d(out)))
        return out

    def scf(self, threshold: float = 5e-3, max_steps: int = 100, policy: str = "fp64",
            switch_ev: float = 1e-3):
        """SCF for every molecule; policy "fp64" (the reference), "mixed" (FP32
        Fock and density, FP64 eigensolver), "mixed_bf16" (FP32 Fock, split-BF16
        density) or "mixed_eig32" (also the eigensolver's reduction and
        back-transform in FP32; opt-in)."""
        res = (ScfResult * self.n_mol)()
        pol = POLICIES[policy]
        _check(self.L.sg_scf_run_policy(self.h, threshold, max_steps, pol, switch_ev, res))
        return [{"e_el": r.e_el, "e_cc": r.e_cc, "e_tot": r.e_tot, "steps": r.steps,
                 "converged": bool(r.converged), "status": r.status} for r in res]

    def set_damping(self, beta: float):
        """Density damping of later scf() calls: P <- beta P_old + (1 - beta) P_new
        each step (0 = off, the reference's SCF)."""
        _check(self.L.sg_scf_set_damping(self.h, float(beta)))

    def set_diis(self, depth: int):
        """Pulay DIIS of later scf() calls over the last `depth` Fock matrices
        (0 = off, the reference's SCF; else 2..8)."""
        _check(self.L.sg_scf_set_diis(self.h, int(depth)))

    def gradient(self) -> np.ndarray:
        """Analytic nuclear gradient (eV/A) of every atom of the batch at the
        density of the last scf() (sg_gradient): an (atoms, 3) array in the
        order the molecules were given."""
        out = np.zeros(3 * int(self.offsets[-1]))
        _check(self.L.sg_gradient(self.h, _d(out)))
        return out.reshape(-1, 3)

    def record_energies(self, cap: int = 101):
        """Record every molecule's E_el per SCF iteration in later scf() calls
        (sg_scf_record_energies): the guess and steps 1 .. cap - 1; 0 stops."""
        self.cap = int(cap)
        _check(self.L.sg_scf_record_energies(self.h, self.cap))

    def energy_history(self, mol: int) -> np.ndarray:
        """E_el (eV) of molecule mol after the guess and each step of the last
        scf(): an array of min(steps + 1, cap) entries."""
        out = np.zeros(max(1, getattr(self, "cap", 0)))
        n = ctypes.c_int(0)
        _check(self.L.sg_scf_energy_history(self.h, int(mol), _d(out), ctypes.byref(n)))
        return out[: n.value].copy()

    def fp32_steps(self) -> np.ndarray:
        """FP32 steps of each molecule in the last scf() (0 for policy fp64)."""
        out = np.zeros(self.n_mol, dtype=np.int32)
        _check(self.L.sg_scf_fp32_steps(self.h, out.ctypes.data_as(_ip)))
        return out

    # host views ---------------------------------------------------------
    def pair_blocks(self, m: int) -> np.ndarray:
        n = self.L.sg_pair_block_count(self.h, m)
        out = np.zeros(n)
        _check(self.L.sg_get_pair_blocks(self.h, m, _d(out)))
        return out

    def packed(self, m: int, which: int) -> np.ndarray:
        n = self.n_ao(m)
        out = np.zeros(n * (n + 1) // 2)
        _check(self.L.sg_get_packed(self.h, m, which, _d(out)))
        return out

    def set_packed(self, m: int, which: int, values: np.ndarray):
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(self.L.sg_set_packed(self.h, m, which, _d(v)))

    def eigen(self, m: int):
        n = self.n_ao(m)
        w = np.zeros(n)
        c = np.zeros(n * n)
        _check(self.L.sg_get_eigen(self.h, m, _d(w), _d(c)))
        return w, c.reshape(n, n).T.copy()  # columns are eigenvectors


class Server:
    """Continuous-batching SCF server (sg_server): submit molecules, step, poll."""

    def __init__(self, engine: Engine, hbm_bytes: float = 8e9, max_slots: int = 4096,
                 max_atoms: int = 1 << 18, threshold: float = 5e-3, max_steps: int = 100,
                 policy: str = "fp64", switch_ev: float = 1e-3, classes: Sequence[int] = ()):
        """classes: increasing basis-size bounds of the size classes (empty: one class)."""
        self.engine = engine
        self.L = engine.L
        h = _vp()
        pol = POLICIES[policy]
        bounds = np.asarray(list(classes), dtype=np.int32)
        _check(self.L.sg_server_create_classes(engine.h, float(hbm_bytes), max_slots, max_atoms,
                                               threshold, max_steps, pol, switch_ev,
                                               len(bounds) + 1, _i(bounds), ctypes.byref(h)))
        self.h = h
        engine._servers.add(self)

    def close(self):
        if self.h:
            self.L.sg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_diis(self, depth: int):
        """Pulay DIIS of the server's SCFs (0 = off, else 2..8); only while idle."""
        _check(self.L.sg_server_set_diis(self.h, int(depth)))

    def submit(self, request_id: int, molecule) -> None:
        Z = np.asarray([int(a[0]) for a in molecule], dtype=np.int32)
        xyz = np.asarray([[float(c) for c in a[1:4]] for a in molecule], dtype=np.float64).ravel()
        _check(self.L.sg_server_submit(self.h, request_id, len(Z), _i(Z), _d(xyz)))

    def step(self):
        """One retirement + admission + iteration round of every ready size class
        -> (in flight, queued)."""
        n_in, n_q = ctypes.c_int(), ctypes.c_int()
        _check(self.L.sg_server_step(self.h, ctypes.byref(n_in), ctypes.byref(n_q)))
        return n_in.value, n_q.value

    def poll(self, max_out: int = 4096):
        buf = (Served * max_out)()
        n = self.L.sg_server_poll(self.h, buf, max_out)
        if n < 0:
            _check(n)
        return [{"id": r.request_id, "e_el": r.e_el, "e_cc": r.e_cc, "e_tot": r.e_tot,
                 "steps": r.steps, "converged": bool(r.converged), "status": r.status,
                 "n_ao": r.n_ao, "t_submit": r.t_submit, "t_admit": r.t_admit,
                 "t_done": r.t_done} for r in buf[:n]]

    def clock(self) -> float:
        return self.L.sg_server_clock(self.h)

    def drain(self, max_seconds: float = 600.0):
        """Steps until nothing is queued or in flight; returns every result."""
        out = []
        t_end = self.clock() + max_seconds
        while self.clock() < t_end:
            n_in, n_q = self.step()
            out.extend(self.poll())
            if n_in == 0 and n_q == 0:
                break
        out.extend(self.poll())
        return out
