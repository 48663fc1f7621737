"""The build at the size of the benchmark, all-trans C20H42 (nbf 490, naux
2366, nocc 81): many exchange tiles, the widest integer column group and every
Q-split in use. Integer and FP64 builds are deterministic run to run (bitwise;
this catches unordered shared-memory and cross-stream accesses), and the
integer build stays within its error bound of FP64."""
import sys

import numpy as np
import pytest

from nopt_amd.molecule import Molecule
from nopt_amd.jk import JKContext
from nopt_amd import geometry as G

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c20():
    m = Molecule.from_angstrom(*G.alkane(20))
    ctx = JKContext(m.nbf, m.naux)
    ctx.set_B(m.eri3_packed(), m.metric())
    rng = np.random.default_rng(7)
    C, _ = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))
    Co = np.ascontiguousarray(C[:, :m.nocc].T) * 0.5
    D = Co.T @ Co
    yield m, ctx, D, Co
    ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_builds_are_bitwise_reproducible(c20, prec):
    m, ctx, D, Co = c20
    say("C20H42", prec, "nbf", m.nbf, "naux", m.naux, "nocc", m.nocc)
    J0, K0 = ctx.build(D, Co, m.nocc, prec=prec)
    for rep in range(4):
        J, K = ctx.build(D, Co, m.nocc, prec=prec)
        dj, dk = np.abs(J - J0).max(), np.abs(K - K0).max()
        say("  rep", rep, "max|J - J0|", dj, "max|K - K0|", dk)
        assert dj == 0.0 and dk == 0.0


def test_int8_within_bound_of_fp64(c20):
    m, ctx, D, Co = c20
    J64, K64 = ctx.build(D, Co, m.nocc, prec="fp64")
    J, K = ctx.build(D, Co, m.nocc, prec="int8")
    dE = np.sum(D * (2 * (J - J64) - (K - K64)))
    say("int8 vs fp64: max|dJ|", np.abs(J - J64).max(), "max|dK|", np.abs(K - K64).max(), "dE", dE)
    assert np.abs(J - J64).max() < 1e-8 * max(1.0, np.abs(J64).max())
    assert np.abs(K - K64).max() < 1e-7 * max(1.0, np.abs(K64).max())
    assert abs(dE) < 1e-6


def test_adjoint_200_rotations(c20):
    """The energy adjoint at the benchmark size: the converged FP64 orbitals
    rotated by exp(kappa), kappa occupied-virtual with entries N(0, 1e-2),
    200 seeds. For every seed the FP64 adjoint dE/dC_occ = 4 F C_occ matches
    central differences (h = 1e-5) of the FP64 energy along a random unit
    direction, the integer-mode energy stays within 1 uEh of FP64, and the
    adjoints of the integer, FP32 and bf16x6 modes stay within their error
    bounds of the FP64 adjoint."""
    from nopt_amd.scf import rhf, rotate
    m, ctx, _, _ = c20
    S, T, V = m.one_electron()
    H = T + V
    nocc, n = m.nocc, m.nbf
    C = rhf(m, ctx=ctx, H=H, S=S, e_conv=1e-10, d_conv=1e-8).C
    precs = ("int8", "fp32", "bf16x6")
    tol = {"int8": 5e-6, "fp32": 5e-5, "bf16x6": 1e-4}
    worst_fd, worst_g, worst_e = 0.0, {p: 0.0 for p in precs}, {p: 0.0 for p in precs}
    for seed in range(200):
        rng = np.random.default_rng(seed)
        Cr = rotate(C, nocc, 1e-2 * rng.standard_normal((n - nocc, nocc)))
        Co = np.ascontiguousarray(Cr[:, :nocc].T)
        E64, G64 = ctx.energy_gradient(H, Co, nocc, prec="fp64")
        for p in precs:
            Ep, Gp = ctx.energy_gradient(H, Co, nocc, prec=p)
            worst_g[p] = max(worst_g[p], float(np.abs(Gp - G64).max()))
            worst_e[p] = max(worst_e[p], abs(Ep - E64))
        dirn = rng.standard_normal(Co.shape)
        dirn /= np.linalg.norm(dirn)
        h = 1e-5
        Ep = ctx.energy_gradient(H, Co + h * dirn, nocc)[0]
        Em = ctx.energy_gradient(H, Co - h * dirn, nocc)[0]
        fd = (Ep - Em) / (2 * h)
        an = float(np.sum(G64 * dirn))
        worst_fd = max(worst_fd, abs(fd - an) / max(1.0, abs(an)))
        if seed % 50 == 0:
            say("  seed", seed, "fd", fd, "<G,dir>", an)
    say("200 rotations: max rel |FD - <G,dir>|", worst_fd, "max|G - G64|", worst_g, "max|E - E64|", worst_e)
    assert worst_fd < 1e-6
    assert worst_e["int8"] < 1e-6
    for p in precs:
        assert worst_g[p] < tol[p], p


def test_int8_batch_beyond_one_column_group(c20):
    """Two densities stacked (162 orbital columns): the integer half-transform
    takes several column groups, writes X in FP64 and splits it per density
    afterwards (the path of systems with more than 96 occupied orbitals);
    each density's J and K stay within the bound of its own FP64 build."""
    m, ctx, D, Co = c20
    C2 = np.ascontiguousarray(Co[::-1] * 0.8)
    D2 = C2.T @ C2
    out = ctx.build_batch([(D, Co, m.nocc), (D2, C2, m.nocc)], prec="int8")
    for (J, K), (Dd, Cd) in zip(out, ((D, Co), (D2, C2))):
        J64, K64 = ctx.build(Dd, Cd, m.nocc, prec="fp64")
        dE = np.sum(Dd * (2 * (J - J64) - (K - K64)))
        say("batch of two, int8 vs fp64: max|dJ|", np.abs(J - J64).max(), "max|dK|", np.abs(K - K64).max(), "dE", dE)
        assert np.abs(J - J64).max() < 1e-8 * max(1.0, np.abs(J64).max())
        assert np.abs(K - K64).max() < 1e-7 * max(1.0, np.abs(K64).max())
        assert abs(dE) < 1e-6


def test_int8_four_column_tiles():
    """C8H18 (33 occupied orbitals): the integer half-transform with four
    16-column tiles (33..64 orbital columns, pair rows) against FP64."""
    m = Molecule.from_angstrom(*G.alkane(8))
    assert 32 < m.nocc <= 64
    with JKContext(m.nbf, m.naux) as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        rng = np.random.default_rng(3)
        C, _ = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))
        Co = np.ascontiguousarray(C[:, :m.nocc].T)
        D = Co.T @ Co
        J64, K64 = ctx.build(D, Co, m.nocc, prec="fp64")
        J, K = ctx.build(D, Co, m.nocc, prec="int8")
        dE = np.sum(D * (2 * (J - J64) - (K - K64)))
        say("C8H18 int8 vs fp64: max|dJ|", np.abs(J - J64).max(), "max|dK|", np.abs(K - K64).max(), "dE", dE)
        assert np.abs(J - J64).max() < 1e-8 * max(1.0, np.abs(J64).max())
        assert np.abs(K - K64).max() < 1e-7 * max(1.0, np.abs(K64).max())
        assert abs(dE) < 1e-6
