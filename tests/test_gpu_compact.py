"""The compact integer form of B and the per-column-group split of the
exchange operand.

A compact context (NOPT_CTX_COMPACT, and every context whose full-row digit
planes would pass 24 GiB) keeps B's lower triangle alone; the half-transform
reads an element above the diagonal from its mirror, stored at the scale of
the other row, and brings it to its own row's scale. The Coulomb pair reads
the same triangle in either context, so J is bitwise the same.

More occupied orbitals than one column group holds (96) cut the exchange
operand into one segment per column group, each with its own row exponents;
more segments than the fused split takes (8) fall back to FP64 X and a
separate split. Every case is checked against the FP64 build of the same
context (itself checked against the oracle in test_gpu_configs.py), and the
NH3 case against the oracle directly."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from nopt_amd import geometry as G
from nopt_amd.jk import JKContext
from nopt_amd.molecule import Molecule
from nopt_amd.workloads import two_electron_energy

pytestmark = pytest.mark.gpu
BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nopt_amd", "basis")
NH3 = (["N", "H", "H", "H"], [[0, 0, 0], [0.94, 0, 0.38], [-0.47, 0.81, 0.38], [-0.47, -0.81, 0.38]])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def core_orbitals(m):
    """Core-Hamiltonian orbitals as rows (all nbf of them)."""
    S, T, V = m.one_electron()
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    C = X @ np.linalg.eigh(X.T @ (T + V) @ X)[1]
    return np.ascontiguousarray(C.T)


def test_nh3_compact_against_oracle():
    """Odd auxiliary count (the last function pair has one member), Cartesian
    orbitals: the compact int8 build against the oracle, J bitwise equal to
    the full-row context's, K within the integer bound of both."""
    om = O.OracleMolecule.from_angstrom(NH3[0], NH3[1], os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-jkfit.gbs"))
    r = O.rhf(om)
    pm = Molecule.from_angstrom(*NH3, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    assert pm.naux % 2 == 1
    C = r["C"].T
    J0, K0 = r["jk"](r["D"], C, om.nocc)
    out = {}
    for compact in (False, True):
        with JKContext(pm.nbf, pm.naux, keep="int8", compact=compact) as ctx:
            ctx.set_B(pm.eri3_packed(), pm.metric())
            out[compact] = ctx.build(r["D"], C, om.nocc, prec="int8") + (ctx.device_bytes(),)
    (Jf, Kf, bf), (Jc, Kc, bc) = out[False], out[True]
    tol = 5e-8 * max(1.0, np.abs(K0).max())
    say("NH3 compact: max|dJ|", np.abs(Jc - J0).max(), "max|dK|", np.abs(Kc - K0).max(),
        "full-row max|dK|", np.abs(Kf - K0).max(), "bytes", bc, "vs", bf)
    assert np.array_equal(Jc, Jf)
    assert np.abs(Jc - J0).max() < tol and np.abs(Kc - K0).max() < tol
    assert bf - bc == 4 * pm.naux * pm.nbf * ((pm.nbf + 63) // 64 * 64)


def test_h2_compact_smallest_shape():
    """10 basis functions (one partial 16-column segment, one partial tile),
    46 auxiliary functions, one occupied orbital."""
    sym, xyz = ["H", "H"], [[0, 0, 0], [0, 0, 0.74]]
    om = O.OracleMolecule.from_angstrom(sym, xyz, os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-jkfit.gbs"))
    r = O.rhf(om)
    pm = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    C = r["C"].T[:1]
    J0, K0 = r["jk"](r["D"], C, 1)
    with JKContext(pm.nbf, pm.naux, compact=True) as ctx:
        ctx.set_B(pm.eri3_packed(), pm.metric())
        J, K = ctx.build(r["D"], C, 1, prec="int8")
    dj, dk = np.abs(J - J0).max(), np.abs(K - K0).max()
    say("H2 compact: max|dJ|", dj, "max|dK|", dk)
    assert dj < 5e-8 * max(1.0, np.abs(J0).max()) and dk < 5e-8 * max(1.0, np.abs(K0).max())


@pytest.fixture(scope="module")
def c20():
    sym, xyz = G.alkane(20)
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    assert (m.nbf, m.naux, m.nocc) == (490, 2366, 81)
    eri3, metric = m.eri3_packed(), m.metric()
    full = JKContext(m.nbf, m.naux)
    full.set_B(eri3, metric)
    compact = JKContext(m.nbf, m.naux, keep="int8", compact=True)
    compact.set_B(eri3, metric)
    yield m, full, compact, core_orbitals(m)
    full.close()
    compact.close()


@pytest.mark.parametrize("nocc", [81, 150, 490])
def test_c20_column_groups(c20, nocc):
    """81 orbitals: one column group; 150: two (96 + 54 columns, a segment
    each); 490: six. Both B forms against the FP64 build."""
    m, full, compact, Call = c20
    C = np.ascontiguousarray(Call[:nocc])
    D = C.T @ C
    J64, K64 = full.build(D, C, nocc, prec="fp64")
    e64 = two_electron_energy(D, J64, K64)
    res = {}
    for name, ctx in (("full rows", full), ("compact", compact)):
        J, K = ctx.build(D, C, nocc, prec="int8")
        dE = two_electron_energy(D, J, K) - e64
        dk = float(np.abs(K - K64).max())
        say("C20H42 nocc", nocc, name, "max|dJ|", float(np.abs(J - J64).max()), "max|dK|", dk, "dE", dE)
        assert dk < 2e-7 * max(1.0, float(np.abs(K64).max())) and abs(dE) < 1e-6 * max(1.0, nocc / 81)
        res[name] = (J, K)
    assert np.array_equal(res["full rows"][0], res["compact"][0])


def test_c20_batch_beyond_the_fused_split(c20):
    """Four 200-orbital factors in one batch: 16 column-group segments, more
    than the fused split takes, so X goes through FP64 and a separate split;
    each K against its single FP64 build."""
    m, full, compact, Call = c20
    rng = np.random.default_rng(3)
    items, refs = [], []
    for k in range(4):
        C = np.ascontiguousarray(Call[rng.permutation(m.nbf)[:200]])
        D = C.T @ C
        items.append((D, C, 200))
        refs.append(full.build(D, C, 200, prec="fp64"))
    for name, ctx in (("full rows", full), ("compact", compact)):
        out = ctx.build_batch(items, prec="int8")
        for (J, K), (J64, K64), (D, C, _) in zip(out, refs, items):
            dk = float(np.abs(K - K64).max())
            dE = two_electron_energy(D, J, K) - two_electron_energy(D, J64, K64)
            say("batch of four 200-orbital factors,", name, "max|dK|", dk, "dE", dE)
            assert dk < 2e-7 * max(1.0, float(np.abs(K64).max())) and abs(dE) < 3e-6
