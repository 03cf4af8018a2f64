"""sg_eigh_noalloc (the eigendecomposition without allocation) against sg_eigh.

Both run the same kernels on the same matrix, so eigenvalues and eigenvectors
must agree bit for bit.  The Fock matrices come from the library's own guess
Fock build of each molecule; the caller's device buffers for the
no-allocation call are allocated here through the HIP runtime the library is
linked against (libamdhip64, test plumbing only) and passed as plain device
pointers through the C ABI.  Two calls in a row, the second on a smaller
matrix, then the size limit.
"""
import ctypes
import sys

import numpy as np
import pytest

from semigpu_amd import api, geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def guess_fock_and_eigh(engine, mol):
    """The guess Fock matrix (full, row-major) and sg_eigh's eigenpairs of it."""
    b = engine.batch([mol])
    b.two_center()
    b.hcore()
    b.guess_density()
    b.fock()
    n = b.n_ao(0)
    f = b.packed(0, api.SG_FOCK)
    full = np.zeros((n, n))
    r, c = np.tril_indices(n)
    full[r, c] = f
    full[c, r] = f
    b.eigh()
    w, vecs = b.eigen(0)
    b.close()
    return full, w, vecs


class Hip:
    """hipMalloc / hipMemcpy / hipStreamCreate of the runtime libsemigpu.so uses."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]

    def ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.ok(self.L.hipMalloc(ctypes.byref(p), nbytes))
        return p.value

    def upload(self, ptr, arr):
        arr = np.ascontiguousarray(arr)
        self.ok(self.L.hipMemcpy(ptr, arr.ctypes.data, arr.nbytes, self.H2D))

    def download(self, ptr, n):
        out = np.empty(n)
        self.ok(self.L.hipMemcpy(out.ctypes.data, ptr, out.nbytes, self.D2H))
        return out


def test_eigh_noalloc_equals_eigh(engine):
    hip = Hip()
    stream = ctypes.c_void_p()
    hip.ok(hip.L.hipStreamCreate(ctypes.byref(stream)))
    cases = [("(H2O)80", geometry.water_cluster(80, geometry.SplitMix64(80)), True),  # N = 480
             ("pyridine", geometry.pyridine(), False)]                               # N = 29
    for name, mol, own_stream in cases:
        F, w_ref, c_ref = guess_fock_and_eigh(engine, mol)
        n = F.shape[0]
        say("case sg_eigh_noalloc", name, "N", n, "caller's stream" if own_stream else "context stream")
        a, w, z = hip.malloc(8 * n * n), hip.malloc(8 * n), hip.malloc(8 * n * n)
        hip.upload(a, F)
        hip.upload(w, np.full(n, np.nan))
        hip.upload(z, np.full(n * n, np.nan))
        engine.eigh_noalloc(n, a, w, z, stream.value if own_stream else 0)
        if own_stream:
            hip.ok(hip.L.hipStreamSynchronize(stream))
        else:  # the context's stream: wait for the whole device
            hip.ok(hip.L.hipDeviceSynchronize())
        W = hip.download(w, n)
        Z = hip.download(z, n * n).reshape(n, n).T  # column k = eigenvector k
        for p in (a, w, z):
            hip.ok(hip.L.hipFree(p))
        say(name, "max |dw|", float(np.abs(W - w_ref).max()), "max |dC|",
            float(np.abs(Z - c_ref).max()))
        assert np.array_equal(W, w_ref), name
        assert np.array_equal(Z, c_ref), name
        # and it is an eigendecomposition of F
        assert np.abs(F @ Z - Z * W).max() <= 1e-9 * np.abs(F).max() * n
    hip.ok(hip.L.hipStreamDestroy(stream))


def test_eigh_noalloc_rejects_beyond_limit(engine):
    with pytest.raises(api.SemiGpuError, match="outside 1..512"):
        engine.eigh_noalloc(513, 1, 1, 1)
    with pytest.raises(api.SemiGpuError, match="outside 1..512"):
        engine.eigh_noalloc(0, 1, 1, 1)
