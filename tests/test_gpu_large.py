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
