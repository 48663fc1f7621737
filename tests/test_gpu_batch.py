"""GPU parity of the batched build (several densities against one B):
the CASSCF pair of builds (src/CAS.cpp:696-707: inactive density and total
density with the factor [C_core; NO sqrt(n/2)], src/molecule.cpp:2959-2977)
and a signed factor of an indefinite perturbed density, against the oracle."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from nopt_amd.molecule import Molecule
from nopt_amd.jk import JKContext
from systems import SYSTEMS

pytestmark = pytest.mark.gpu
BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nopt_amd", "basis")
TOL = {"fp64": 1e-10, "int8": 5e-8}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c2h4():
    sym, xyz = SYSTEMS["C2H4"]
    om = O.OracleMolecule.from_angstrom(sym, xyz, os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-ri.gbs"))
    r = O.rhf(om)
    pm = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-ri", pure=False)
    ctx = JKContext(pm.nbf, pm.naux)
    ctx.set_B(pm.eri3_packed(), pm.metric())
    yield om, r, pm, ctx
    ctx.close()


def cas_pair(r, ncore, nact, seed=0):
    """Inactive (D_core, C_core) and total (2 D_core + D_act, [C_core; NO sqrt(n/2)])
    with a random PSD active 1-RDM of trace nact (natural orbitals NO)."""
    C = r["C"].T                                   # orbitals as rows
    rng = np.random.default_rng(seed)
    Cc, Ca = C[:ncore], C[ncore:ncore + nact]
    A = rng.standard_normal((nact, nact))
    G = A @ A.T
    G *= nact / np.trace(G)
    occ, U = np.linalg.eigh(G)
    occ = np.clip(occ, 0.0, None)
    NO = U.T @ Ca
    Dc = Cc.T @ Cc
    Dt = 2.0 * Dc + 0.5 * NO.T @ np.diag(occ) @ NO
    Ct = np.vstack([Cc * np.sqrt(2.0), NO * np.sqrt(0.5 * occ)[:, None]])
    return (Dc, Cc, ncore), (Dt, Ct, Ct.shape[0])


@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_cas_pair_matches_oracle_and_single_builds(c2h4, prec):
    om, r, pm, ctx = c2h4
    say("C2H4 CAS pair", prec, "n_ao", pm.nbf)
    items = cas_pair(r, ncore=5, nact=3)
    out = ctx.build_batch(list(items), prec=prec)
    for (D, C, no), (J, K) in zip(items, out):
        J0, K0 = r["jk"](D, C, no)
        Js, Ks = ctx.build(D, C, no, prec=prec)
        say("  nocc", no, "max|dJ|", np.abs(J - J0).max(), "max|dK|", np.abs(K - K0).max(),
            "vs single", np.abs(J - Js).max(), np.abs(K - Ks).max())
        # int8: relative to the largest element (the total-density factor carries sqrt(2) C_core)
        assert np.abs(J - J0).max() < TOL[prec] * max(1.0, np.abs(J0).max())
        assert np.abs(K - K0).max() < TOL[prec] * max(1.0, np.abs(K0).max())
        assert np.abs(J - Js).max() < 1e-12 and np.abs(K - Ks).max() < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_signed_factor_of_perturbed_density(c2h4, prec):
    """D_k = D + eps sym(R) is indefinite: K = K(X+) - K(X-) from its signed
    eigen-factor, J from D_k; checked against the explicit four-index RI J/K."""
    om, r, pm, ctx = c2h4
    n = pm.nbf
    rng = np.random.default_rng(0)
    R = rng.standard_normal((n, n))
    Dk = r["D"] + 1e-3 * 0.5 * (R + R.T)
    s, X = np.linalg.eigh(Dk)
    Xp = (X[:, s > 0] * np.sqrt(s[s > 0])).T
    Xm = (X[:, s < 0] * np.sqrt(-s[s < 0])).T
    say("signed factor", prec, "rank+", Xp.shape[0], "rank-", Xm.shape[0])
    (J, Kp), (_, Km) = ctx.build_batch([(Dk, Xp, Xp.shape[0]), (None, Xm, Xm.shape[0])], prec=prec)
    Jb, Kb = O.brute_force_jk(om.eri3_nopt(), om.metric(), Dk)
    K = Kp - Km
    say("  max|dJ|", np.abs(J - Jb).max(), "max|dK|", np.abs(K - Kb).max())
    assert np.abs(J - Jb).max() < (1e-9 if prec == "fp64" else 5e-8)
    assert np.abs(K - Kb).max() < (1e-9 if prec == "fp64" else 1e-7)


def test_batch_of_four_and_empty_entries(c2h4):
    """Four densities, one with no occupied orbitals (K = 0) and one J-only."""
    om, r, pm, ctx = c2h4
    C = r["C"].T
    D = r["D"]
    items = [(D, C, om.nocc), (None, C[:3], 3), (D, C, 0), (0.5 * D, None, 0)]
    out = ctx.build_batch(items, prec="fp64")
    J0, K0 = r["jk"](D, C, om.nocc)
    _, K3 = r["jk"](D, C[:3], 3)
    assert np.abs(out[0][0] - J0).max() < 1e-10 and np.abs(out[0][1] - K0).max() < 1e-10
    assert out[1][0] is None and np.abs(out[1][1] - K3).max() < 1e-10
    assert np.abs(out[2][0] - J0).max() < 1e-10 and np.abs(out[2][1]).max() == 0.0
    assert np.abs(out[3][0] - 0.5 * J0).max() < 1e-10 and np.abs(out[3][1]).max() == 0.0
