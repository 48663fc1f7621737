"""Densities of the CASSCF-style ragged batch of water clusters.

Each cluster (H2O)_n, in cc-pVDZ / cc-pVDZ-JKFIT, gets the two densities of
one CASSCF macro-iteration (src/CAS.cpp:696-707): an inactive closed-shell
density over all but 5 doubly occupied core-guess orbitals, and a total
density with an active space of 10 orbitals and a random PSD 1-RDM of trace
10, whose K factor is [sqrt(2) C_core; NO sqrt(n/2)] with the natural
orbitals NO and occupations n of the 1-RDM (negative occupations set to zero,
src/molecule.cpp:2959-2977).
"""
import numpy as np

from . import geometry as G
from .molecule import Molecule

N_ACTIVE = 10
N_ACTIVE_OCC = 5  # doubly occupied orbitals moved from the inactive set into the active space


def cluster_densities(m, seed, H=None, S=None):
    """(Dc, Cc, Dt, Ct) for molecule m: inactive density Dc = Cc^T Cc (orbitals
    as rows of Cc) and total density Dt = Ct^T Ct, from the core-Hamiltonian
    orbitals and a 1-RDM drawn from default_rng(seed)."""
    if S is None or H is None:
        S_, T, V = m.one_electron()
        S = S_ if S is None else S
        H = (T + V) if H is None else H
    s_, U = np.linalg.eigh(S)
    X = U @ np.diag(s_ ** -0.5) @ U.T
    C = X @ np.linalg.eigh(X.T @ H @ X)[1]
    ncore = m.nocc - N_ACTIVE_OCC
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((N_ACTIVE, N_ACTIVE))
    gam = A @ A.T
    gam *= float(N_ACTIVE) / np.trace(gam)
    occ, W = np.linalg.eigh(gam)
    occ = np.clip(occ, 0.0, None)
    Cc = np.ascontiguousarray(C[:, :ncore].T)
    NO = W.T @ C[:, ncore:ncore + N_ACTIVE].T
    Ct = np.ascontiguousarray(np.vstack([np.sqrt(2.0) * Cc, NO * np.sqrt(0.5 * occ)[:, None]]))
    return Cc.T @ Cc, Cc, Ct.T @ Ct, Ct


def ragged_clusters(count=32, nmax=16, seed=0, pure=True):
    """The ragged batch: `count` clusters (H2O)_n with n ~ U{1..nmax} (seed),
    cluster k placed with seed + 1 + k and its 1-RDM drawn with the same seed.
    Yields (n, molecule, (Dc, Cc, Dt, Ct)); pure=False gives NOPT's Cartesian
    orbital basis."""
    for k, n in enumerate(G.batch_sizes(count, nmax, seed=seed)):
        sym, xyz = G.water_cluster(n, seed=seed + 1 + k)
        m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=pure)
        yield n, m, cluster_densities(m, seed + 1 + k)


def two_electron_energy(D, J, K, total=False):
    """Two-electron energy of a density: sum D.(2J - K) for an inactive
    (closed-shell, unit occupations) density, a quarter of it for a
    spin-summed total density (occupations 0..2)."""
    e = float(np.sum(D * (2.0 * J - K)))
    return 0.25 * e if total else e
