"""Numpy restatement of the library's opt-in Pulay DIIS SCF over the oracle's
Fock builds (test infrastructure).

The SCF is the reference's (scfcalculation.cpp:82-146: eigensolve, density,
Fock, energy, converged when |E_k - E_(k-1)| < threshold), with one step
added after the convergence test of a molecule that goes on iterating:
store F and e = F P - P F in a ring of `depth` slots, solve
[B -1; -1 0] [c; l] = [0; -1] with B_ij = <e_i, e_j> (scaled by its largest
diagonal entry), and hand sum_j c_j F_j to the next eigensolve.  Pulay,
Chem. Phys. Lett. 73, 393 (1980); J. Comput. Chem. 3, 556 (1982).
"""
import numpy as np

import pyoracle as po


def _pack(a: np.ndarray) -> np.ndarray:
    r, c = np.tril_indices(a.shape[0])
    order = np.lexsort((c, r))
    return a[r[order], c[order]].copy()


def diis_scf(atoms, depth: int, threshold: float, max_steps: int = 100):
    """Returns (e_tot, steps, converged, energies) of the DIIS SCF."""
    om = po.Molecule(atoms)
    n, no = om.n_ao, om.n_occ
    blocks = om.pair_blocks()
    h = om.hcore(blocks)
    e_cc = om.core_core(blocks)
    p = om.guess_density()
    f = om.fock(blocks, h, p)
    e = po.energy(n, h, f, p)
    energies = [e]
    f_eig = po.packed_to_full(f, n)
    hist_f, hist_e = [], []
    steps, conv = 0, False
    while steps < max_steps:
        w, v = np.linalg.eigh(f_eig)
        c = v[:, :no]
        p = _pack(2.0 * c @ c.T)
        f = om.fock(blocks, h, p)
        e_new = po.energy(n, h, f, p)
        steps += 1
        conv = abs(e - e_new) < threshold
        e = e_new
        energies.append(e)
        if conv:
            break
        ff, pf = po.packed_to_full(f, n), po.packed_to_full(p, n)
        hist_f.append(ff)
        hist_e.append(ff @ pf - pf @ ff)
        hist_f, hist_e = hist_f[-depth:], hist_e[-depth:]
        k = len(hist_f)
        f_eig = ff
        if k > 1:
            B = np.array([[np.sum(a * b) for b in hist_e] for a in hist_e])
            scale = B.diagonal().max()
            if scale > 0 and np.isfinite(scale):
                A = np.zeros((k + 1, k + 1))
                A[:k, :k] = B / scale
                A[:k, k] = A[k, :k] = -1.0
                rhs = np.zeros(k + 1)
                rhs[k] = -1.0
                try:
                    x = np.linalg.solve(A, rhs)
                    if np.all(np.isfinite(x[:k])):
                        f_eig = sum(ci * fi for ci, fi in zip(x[:k], hist_f))
                except np.linalg.LinAlgError:
                    pass
    return e + e_cc, steps, conv, np.array(energies)
