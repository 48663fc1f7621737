"""Geometries of the benchmark systems (Angstrom in, bohr out via Molecule).

* water: O at the origin, O-H 0.9572 A, H-O-H 104.52 deg.
* all-trans alkane CnH(2n+2): planar zig-zag carbon backbone with C-C 1.54 A
  and C-C-C 109.4712 deg (tetrahedral); C-H 1.09 A with tetrahedral H-C-H;
  the CH3 termini are tetrahedral.
* water clusters: rigid monomers with uniformly random orientations placed
  uniformly in a cube sized for 1 g/cm^3, rejecting any placement with an O-O
  distance below 2.7 A.

Random numbers come from numpy's PCG64 generator (bit-reproducible for a
given seed); normal deviates use an explicit Box-Muller transform so they do
not depend on numpy's sampling algorithms.
"""
import math

import numpy as np

TETRA = math.acos(-1.0 / 3.0)  # 109.4712 deg
AVOGADRO = 6.02214076e23
M_H2O = 18.015  # g/mol


def water():
    th = math.radians(104.52)
    r = 0.9572
    sym = ["O", "H", "H"]
    xyz = np.array([[0.0, 0.0, 0.0],
                    [r * math.sin(th / 2), 0.0, r * math.cos(th / 2)],
                    [-r * math.sin(th / 2), 0.0, r * math.cos(th / 2)]])
    return sym, xyz


def alkane(n=20, r_cc=1.54, r_ch=1.09):
    """All-trans CnH(2n+2)."""
    dx = r_cc * math.sin(TETRA / 2)
    dy = r_cc * math.cos(TETRA / 2)
    # carbons k = -1 and n are virtual positions used to place the CH3 hydrogens
    C = {k: np.array([k * dx, (k % 2) * dy, 0.0]) for k in range(-1, n + 1)}
    sym, xyz = [], []
    for k in range(n):
        sym.append("C")
        xyz.append(C[k])
    for k in range(n):
        c = C[k]
        u1 = (C[k - 1] - c) / np.linalg.norm(C[k - 1] - c)
        u2 = (C[k + 1] - c) / np.linalg.norm(C[k + 1] - c)
        bis = -(u1 + u2)
        bis /= np.linalg.norm(bis)
        z = np.array([0.0, 0.0, 1.0])
        for s in (+1.0, -1.0):
            d = bis * math.cos(TETRA / 2) + s * z * math.sin(TETRA / 2)
            sym.append("H")
            xyz.append(c + r_ch * d)
        # terminal carbons: one more H toward the virtual backbone neighbour
        if k == 0:
            sym.append("H")
            xyz.append(c + r_ch * u1)
        if k == n - 1:
            sym.append("H")
            xyz.append(c + r_ch * u2)
    return sym, np.array(xyz)


def _normal(rng, size):
    u1 = rng.random(size)
    u2 = rng.random(size)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def random_rotation(rng):
    """Uniform rotation from a normalised Gaussian quaternion."""
    q = _normal(rng, 4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def cube_side(n_water, density=1.0):
    """Cube side in Angstrom holding n waters at `density` g/cm^3."""
    vol_cm3 = n_water * M_H2O / (AVOGADRO * density)
    return (vol_cm3 ** (1.0 / 3.0)) * 1e8


def water_cluster(n, seed, min_oo=2.7, max_tries=100000):
    rng = np.random.Generator(np.random.PCG64(seed))
    wsym, wxyz = water()
    side = cube_side(n)
    oxygens, sym, xyz = [], [], []
    for _ in range(n):
        for _t in range(max_tries):
            o = rng.random(3) * side
            if all(np.linalg.norm(o - p) >= min_oo for p in oxygens):
                break
        else:
            raise RuntimeError("water_cluster: placement failed")
        R = random_rotation(rng)
        oxygens.append(o)
        for s, r in zip(wsym, wxyz):
            sym.append(s)
            xyz.append(o + R @ r)
    return sym, np.array(xyz)


def batch_sizes(count=32, nmax=16, seed=0):
    """Cluster sizes n ~ U{1..nmax} for the ragged batch."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [int(v) for v in rng.integers(1, nmax + 1, size=count)]
