"""Edge shapes of the integer build: an odd number of auxiliary functions
(NH3 / cc-pVDZ-JKFIT, naux 139: the last pair of functions has one member),
a batch of four densities (one with no occupied orbitals, one J-only, one
K-only) against single builds and the oracle, a single occupied orbital, and
the smallest shapes (H2) in every precision mode."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from nopt_amd.jk import JKContext
from nopt_amd.molecule import Molecule

pytestmark = pytest.mark.gpu
BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nopt_amd", "basis")
NH3 = (["N", "H", "H", "H"], [[0, 0, 0], [0.94, 0, 0.38], [-0.47, 0.81, 0.38], [-0.47, -0.81, 0.38]])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nh3():
    om = O.OracleMolecule.from_angstrom(NH3[0], NH3[1], os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-jkfit.gbs"))
    r = O.rhf(om)
    pm = Molecule.from_angstrom(*NH3, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    assert pm.naux % 2 == 1
    ctx = JKContext(pm.nbf, pm.naux)
    ctx.set_B(pm.eri3_packed(), pm.metric())
    yield om, r, pm, ctx
    ctx.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("int8", 5e-8)])
def test_odd_aux_count(nh3, prec, tol):
    om, r, pm, ctx = nh3
    C = r["C"].T
    J0, K0 = r["jk"](r["D"], C, om.nocc)
    J, K = ctx.build(r["D"], C, om.nocc, prec=prec)
    dj, dk = np.abs(J - J0).max(), np.abs(K - K0).max()
    say("NH3 naux", pm.naux, prec, "max|dJ|", dj, "max|dK|", dk)
    assert dj < tol * max(1.0, np.abs(J0).max()) and dk < tol * max(1.0, np.abs(K0).max())


def test_int8_batch_of_four(nh3):
    om, r, pm, ctx = nh3
    C = r["C"].T
    D = r["D"]
    items = [(D, C, om.nocc), (None, C[:3], 3), (D, C, 0), (0.5 * D, None, 0)]
    out = ctx.build_batch(items, prec="int8")
    J0, K0 = r["jk"](D, C, om.nocc)
    _, K3 = r["jk"](D, C[:3], 3)
    Js, Ks = ctx.build(D, C, om.nocc, prec="int8")
    say("batch of four: max|dJ|", np.abs(out[0][0] - J0).max(), "max|dK|", np.abs(out[0][1] - K0).max(),
        "vs single", np.abs(out[0][0] - Js).max(), np.abs(out[0][1] - Ks).max())
    tol = 5e-8 * max(1.0, np.abs(K0).max())
    assert np.abs(out[0][0] - J0).max() < tol and np.abs(out[0][1] - K0).max() < tol
    assert np.abs(out[0][0] - Js).max() < 1e-12 * np.abs(Js).max() and np.abs(out[0][1] - Ks).max() == 0.0
    assert out[1][0] is None and np.abs(out[1][1] - K3).max() < tol
    assert np.abs(out[2][0] - J0).max() < tol and np.abs(out[2][1]).max() == 0.0
    assert np.abs(out[3][0] - 0.5 * J0).max() < tol and np.abs(out[3][1]).max() == 0.0


@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_single_occupied_orbital(nh3, prec):
    om, r, pm, ctx = nh3
    C = r["C"].T[:1]
    D = C.T @ C
    J0, K0 = r["jk"](D, C, 1)
    J, K = ctx.build(D, C, 1, prec=prec)
    tol = (1e-10 if prec == "fp64" else 5e-8) * max(1.0, np.abs(J0).max())
    say("one orbital", prec, "max|dJ|", np.abs(J - J0).max(), "max|dK|", np.abs(K - K0).max())
    assert np.abs(J - J0).max() < tol and np.abs(K - K0).max() < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("int8", 5e-8), ("fp32", 5e-6), ("bf16x6", 5e-6),
                                      ("bf16x3", 2e-4)])
def test_smallest_shapes(prec, tol):
    """H2 / cc-pVDZ / cc-pVDZ-JKFIT, Cartesian: 10 basis functions (less than
    one 16-column segment and one 64-row tile, every tile mostly padding), 46
    auxiliary functions and one occupied orbital, in every precision mode."""
    sym, xyz = ["H", "H"], [[0, 0, 0], [0, 0, 0.74]]
    om = O.OracleMolecule.from_angstrom(sym, xyz, os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-jkfit.gbs"))
    r = O.rhf(om)
    pm = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    assert (pm.nbf, pm.naux, om.nocc) == (10, 46, 1)
    C = r["C"].T[:1]
    J0, K0 = r["jk"](r["D"], C, 1)
    with JKContext(pm.nbf, pm.naux) as ctx:
        ctx.set_B(pm.eri3_packed(), pm.metric())
        J, K = ctx.build(r["D"], C, 1, prec=prec)
        J2, K2 = ctx.build(r["D"], C, 1, prec=prec)
    dj, dk = np.abs(J - J0).max(), np.abs(K - K0).max()
    say("H2", prec, "max|dJ|", dj, "max|dK|", dk)
    assert dj < tol * max(1.0, np.abs(J0).max()) and dk < tol * max(1.0, np.abs(K0).max())
    assert np.array_equal(J, J2) and np.array_equal(K, K2)
