"""GPU tests of the drivers around the build: mixed-precision RHF refined to
the FP64 fixed point (src/SCF.cpp:78-223), direct orbital minimisation with
the adjoint dE/dC = 4 F C_occ (the role of NOPT's SOSCF engine,
src/converger_2_1.cpp:22-123), and the adjoint at rotated orbitals against
central differences."""
import sys

import numpy as np
import pytest

from nopt_amd.molecule import Molecule
from nopt_amd.jk import JKContext
from nopt_amd import geometry as G
from nopt_amd.scf import rhf, minimize_orbitals, rotate
from systems import SYSTEMS, GOLDEN_E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def core_guess(S, H):
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    _, Cp = np.linalg.eigh(X.T @ H @ X)
    return X @ Cp


@pytest.mark.parametrize("name,prec", [("C2H4", "int8"), ("C2H4", "fp32"), ("guanine", "int8")])
def test_mixed_rhf_refines_to_nopt_energy(name, prec):
    """Mixed-precision builds until the DIIS error is below 1e-5, then FP64:
    the converged energy is NOPT's log energy (C2H4, guanine)."""
    sym, xyz = SYSTEMS[name]
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-ri")
    res = rhf(m, prec=prec, e_conv=1e-11, d_conv=1e-9)
    say(name, prec, "E", res.E, "diff", res.E - GOLDEN_E[name], "builds", res.builds)
    assert res.builds.get(prec, 0) > 0 and res.builds.get("fp64", 0) > 0
    assert abs(res.E - GOLDEN_E[name]) < 3e-10


def test_direct_minimisation_reaches_diis_energy():
    """H2O/cc-pVDZ/cc-pVDZ-JKFIT (24 spherical functions): L-BFGS over
    orbital rotations from the core guess converges to the DIIS energy."""
    sym, xyz = G.water()
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    S, T, V = m.one_electron()
    H = T + V
    with JKContext(m.nbf, m.naux) as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        ref = rhf(m, ctx=ctx, H=H, S=S, e_conv=1e-12, d_conv=1e-10)
        res = minimize_orbitals(m, ctx, H, core_guess(S, H), g_conv=1e-7, e_conv=1e-12)
    say("DIIS", ref.E, "direct", res.E, "diff", res.E - ref.E, "iters", res.iters, "gmax", res.gmax)
    assert res.converged
    assert abs(res.E - ref.E) < 1e-9


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("int8", 5e-7)])
def test_adjoint_at_rotated_orbitals(prec, tol):
    """C_occ from exp(kappa) rotations of the converged C2H4 orbitals (kappa
    occupied-virtual, entries N(0, 1e-2), 20 seeds): the directional
    derivative of the FP64 energy by central differences (h = 1e-5) matches
    <G, dir> of the FP64 adjoint, and the adjoint of `prec` matches FP64."""
    sym, xyz = SYSTEMS["C2H4"]
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-ri")
    S, T, V = m.one_electron()
    H = T + V
    nocc, n = m.nocc, m.nbf
    worst_fd, worst_g = 0.0, 0.0
    with JKContext(n, m.naux) as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        C = rhf(m, ctx=ctx, H=H, S=S).C
        for seed in range(20):
            rng = np.random.default_rng(seed)
            Cr = rotate(C, nocc, 1e-2 * rng.standard_normal((n - nocc, nocc)))
            Co = np.ascontiguousarray(Cr[:, :nocc].T)
            _, G64 = ctx.energy_gradient(H, Co, nocc, prec="fp64")
            _, Gp = ctx.energy_gradient(H, Co, nocc, prec=prec)
            dirn = rng.standard_normal(Co.shape)
            dirn /= np.linalg.norm(dirn)
            h = 1e-5
            Ep = ctx.energy_gradient(H, Co + h * dirn, nocc)[0]
            Em = ctx.energy_gradient(H, Co - h * dirn, nocc)[0]
            fd = (Ep - Em) / (2 * h)
            an = float(np.sum(G64 * dirn))
            worst_fd = max(worst_fd, abs(fd - an) / max(1.0, abs(an)))
            worst_g = max(worst_g, float(np.abs(Gp - G64).max()))
    say(prec, "max rel |FD - <G,dir>|", worst_fd, "max|G - G64|", worst_g)
    assert worst_fd < 1e-7
    assert worst_g < tol
