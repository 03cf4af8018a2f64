"""Synthetic geometries for the MNDO SCF workloads.

Every random draw comes from a project-written SplitMix64 stream with explicit
transforms (53-bit uniform doubles, Box-Muller normals, normalised 4-D Gaussian
quaternions), so a seed names the same geometry on every machine.  Coordinates
are in Angstrom, atoms are (Z, x, y, z) like the XYZ files read by
``ReadXYZFile::GetValuesFromFile`` (src/common/readxyzfile.cpp:27-55).

Workloads
---------
* ``pyridine()``            C5H5N, ring radius 1.39 A, N pulled radially so both
                            C-N bonds are 1.34 A, radial C-H 1.08 A.
* ``alkane(n)``             all-trans C_nH_(2n+2): C-C 1.54 A, C-H 1.09 A,
                            tetrahedral angles (109.47 deg), backbone in the xy plane.
* ``water_cluster(n, rng)`` (H2O)_n: O on the first n sites of a ceil(n^(1/3))^3
                            cubic lattice (lexicographic), spacing 2.9 A, jitter
                            U(-0.15, 0.15) A per coordinate, O-H 0.96 A,
                            H-O-H 104.5 deg, uniform random orientation.
* ``water_batch(count, seed)``   the 4096-cluster batch, n ~ U{2..64}.
* ``water_box(seed)``       1000 H2O on a 10x10x10 lattice, spacing 3.1 A, jitter 0.2 A.
* ``serving_stream(rate, seconds, seed)``  Poisson arrivals of alkanes
                            (n ~ U{1..60}) or water clusters (n ~ U{2..200}), 50/50.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

Atom = Tuple[int, float, float, float]

_MASK64 = (1 << 64) - 1
TETRA_DEG = 109.47


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) with explicit output transforms."""

    def __init__(self, seed: int):
        self.state = seed & _MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def uniform_range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.uniform()

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive) by rejection, no modulo bias."""
        span = hi - lo + 1
        limit = (1 << 64) - ((1 << 64) % span)
        while True:
            x = self.next_u64()
            if x < limit:
                return lo + x % span

    def normal(self) -> float:
        """Standard normal by Box-Muller (one value per call, second discarded)."""
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def exponential(self, rate: float) -> float:
        u = self.uniform()
        return -math.log1p(-u) / rate

    def rotation(self) -> List[List[float]]:
        """Uniform random rotation matrix from a normalised 4-D Gaussian quaternion."""
        while True:
            q = [self.normal() for _ in range(4)]
            nrm = math.sqrt(sum(c * c for c in q))
            if nrm > 1e-12:
                break
        w, x, y, z = (c / nrm for c in q)
        return [
            [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
        ]


def _add(a, b):
    return [a[0] + b[0], a[1] + b[1], a[2] + b[2]]


def _sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def _scale(a, s):
    return [a[0] * s, a[1] * s, a[2] * s]


def _unit(a):
    n = math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    return [a[0] / n, a[1] / n, a[2] / n]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def pyridine() -> List[Atom]:
    """C5H5N: C on a 1.39 A hexagon, N at vertex 0 moved radially to C-N = 1.34 A."""
    r = 1.39
    c1 = (r * math.cos(math.radians(60.0)), r * math.sin(math.radians(60.0)))
    # (x_N - c1x)^2 + c1y^2 = 1.34^2, N on the +x axis outside c1x
    x_n = c1[0] + math.sqrt(1.34 ** 2 - c1[1] ** 2)
    atoms: List[Atom] = [(7, x_n, 0.0, 0.0)]
    for k in range(1, 6):
        th = math.radians(60.0 * k)
        atoms.append((6, r * math.cos(th), r * math.sin(th), 0.0))
    for k in range(1, 6):
        th = math.radians(60.0 * k)
        atoms.append((1, (r + 1.08) * math.cos(th), (r + 1.08) * math.sin(th), 0.0))
    return atoms


def alkane(n: int) -> List[Atom]:
    """All-trans C_nH_(2n+2); carbons first, then hydrogens."""
    if n < 1:
        raise ValueError("alkane needs n >= 1")
    cc, ch = 1.54, 1.09
    half = math.radians(TETRA_DEG / 2.0)
    if n == 1:
        d = ch / math.sqrt(3.0)
        hs = [[d, d, d], [d, -d, -d], [-d, d, -d], [-d, -d, d]]
        return [(6, 0.0, 0.0, 0.0)] + [(1, h[0], h[1], h[2]) for h in hs]
    dx = cc * math.sin(half)
    dy = cc * math.cos(half)
    # carbons k = -1 .. n carry the virtual neighbours of the two chain ends
    pos = {k: [k * dx, (dy / 2.0) if k % 2 == 0 else (-dy / 2.0), 0.0] for k in range(-1, n + 1)}
    zhat = [0.0, 0.0, 1.0]
    carbons = [pos[k] for k in range(n)]
    hydrogens = []
    for k in range(n):
        c = pos[k]
        b1 = _unit(_sub(pos[k - 1], c))
        b2 = _unit(_sub(pos[k + 1], c))
        bis = _unit(_add(b1, b2))
        # two hydrogens out of the backbone plane, tetrahedral
        for sgn in (1.0, -1.0):
            u = _add(_scale(bis, -math.cos(half)), _scale(zhat, sgn * math.sin(half)))
            hydrogens.append(_add(c, _scale(_unit(u), ch)))
        if k == 0:
            hydrogens.append(_add(c, _scale(b1, ch)))
        if k == n - 1:
            hydrogens.append(_add(c, _scale(b2, ch)))
    return [(6, *c) for c in carbons] + [(1, *h) for h in hydrogens]


def _lattice_sites(n: int) -> List[Tuple[int, int, int]]:
    m = 1
    while m * m * m < n:
        m += 1
    out = []
    for i in range(m):
        for j in range(m):
            for k in range(m):
                if len(out) < n:
                    out.append((i, j, k))
    return out


def _water_local() -> List[List[float]]:
    oh = 0.96
    half = math.radians(104.5 / 2.0)
    return [[oh * math.sin(half), 0.0, oh * math.cos(half)],
            [-oh * math.sin(half), 0.0, oh * math.cos(half)]]


def water_cluster(n: int, rng: SplitMix64, spacing: float = 2.9, jitter: float = 0.15,
                  sites: Sequence[Tuple[int, int, int]] = None) -> List[Atom]:
    """(H2O)_n; per water O, H, H."""
    sites = list(sites) if sites is not None else _lattice_sites(n)
    local = _water_local()
    atoms: List[Atom] = []
    for s in sites[:n]:
        o = [s[0] * spacing + rng.uniform_range(-jitter, jitter),
             s[1] * spacing + rng.uniform_range(-jitter, jitter),
             s[2] * spacing + rng.uniform_range(-jitter, jitter)]
        rot = rng.rotation()
        atoms.append((8, o[0], o[1], o[2]))
        for h in local:
            v = [sum(rot[r][c] * h[c] for c in range(3)) for r in range(3)]
            atoms.append((1, o[0] + v[0], o[1] + v[1], o[2] + v[2]))
    return atoms


def water_batch(count: int = 4096, seed: int = 20240607, n_min: int = 2,
                n_max: int = 64) -> List[List[Atom]]:
    rng = SplitMix64(seed)
    sizes = [rng.randint(n_min, n_max) for _ in range(count)]
    return [water_cluster(n, rng) for n in sizes]


def water_box(seed: int = 1000, side: int = 10, spacing: float = 3.1,
              jitter: float = 0.2) -> List[Atom]:
    rng = SplitMix64(seed)
    sites = [(i, j, k) for i in range(side) for j in range(side) for k in range(side)]
    return water_cluster(len(sites), rng, spacing=spacing, jitter=jitter, sites=sites)


def serving_stream(rate: float, seconds: float, seed: int = 7
                   ) -> List[Tuple[float, str, List[Atom]]]:
    """Poisson arrivals: (arrival time s, label, atoms)."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(rate)
        if t >= seconds:
            break
        if rng.uniform() < 0.5:
            n = rng.randint(1, 60)
            out.append((t, "C%dH%d" % (n, 2 * n + 2), alkane(n)))
        else:
            n = rng.randint(2, 200)
            out.append((t, "(H2O)%d" % n, water_cluster(n, rng)))
    return out


def n_basis(atoms: Sequence[Atom]) -> int:
    return sum(1 if a[0] <= 2 else 4 for a in atoms)


def n_electrons(atoms: Sequence[Atom]) -> int:
    """Valence electrons from the MNDO core charges (src/common/densitymatrix.cpp:99-107)."""
    tot = 0
    for a in atoms:
        z = a[0]
        tot += z if z <= 2 else (z - 2 if z <= 10 else z - 10)
    return tot


def to_xyz(atoms: Sequence[Atom], comment: str = "") -> str:
    sym = {1: "H", 6: "C", 7: "N", 8: "O"}
    lines = [str(len(atoms)), comment]
    for z, x, y, w in atoms:
        lines.append("%s %.10f %.10f %.10f" % (sym[z], x, y, w))
    return "\n".join(lines) + "\n"
