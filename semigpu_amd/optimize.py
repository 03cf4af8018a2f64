"""Geometry optimisation on the MNDO surface (README.md:109-110 lists a
geometry optimiser as to do; the reference has none).

Quasi-Newton minimisation (BFGS, scipy.optimize.minimize) of the SCF total
energy over the Cartesian coordinates.  Each evaluation is, by default, one
SCF of the molecule followed by the analytic gradient at its converged
density (sg_gradient); gradient="fd" takes instead one call of
sg_gradient_fd (the molecule and its 6 n_atoms displaced copies in one batched
SCF, central differences).
"""
from __future__ import annotations

from typing import Dict, Sequence

import numpy as np

from . import api


def optimize(engine: api.Engine, molecule: Sequence, gtol: float = 1e-3, max_iter: int = 200,
             step: float = 1e-3, threshold: float = 1e-9, diis: int = 6,
             gradient: str = "analytic") -> Dict[str, object]:
    """Minimises E_tot of molecule [(Z, x, y, z) in Angstrom, ...] until the
    largest gradient component is below gtol (eV/A).  gradient is "analytic"
    (sg_gradient) or "fd" (sg_gradient_fd with the given step).  Returns the
    geometry, E_tot (eV), the gradient, the number of gradient evaluations and
    whether scipy reported success."""
    if gradient not in ("analytic", "fd"):
        raise ValueError("gradient must be 'analytic' or 'fd'")
    from scipy.optimize import minimize

    Z = [int(a[0]) for a in molecule]
    x0 = np.array([a[1:4] for a in molecule], dtype=np.float64).reshape(-1)
    evals = [0]

    def atoms(x):
        r = x.reshape(-1, 3)
        return [(z, float(p[0]), float(p[1]), float(p[2])) for z, p in zip(Z, r)]

    def evaluate(x):
        if gradient == "analytic":
            return engine.gradient(atoms(x), threshold=threshold, diis=diis)
        return engine.gradient_fd(atoms(x), step=step, threshold=threshold, diis=diis)

    def fun(x):
        g, e, conv = evaluate(x)
        if not conv:
            raise api.SemiGpuError("SCF did not converge during the geometry optimisation")
        evals[0] += 1
        return e, g.reshape(-1)

    res = minimize(fun, x0, jac=True, method="BFGS",
                   options={"gtol": gtol, "norm": np.inf, "maxiter": max_iter})
    g, e, _ = evaluate(res.x)
    return {"atoms": atoms(res.x), "e_tot": e, "gradient": g, "evaluations": evals[0],
            "success": bool(res.success)}
