"""Contexts that keep one form of B, formed in several chunks of rows: a
water cluster (H2O)32 in cc-pVDZ / cc-pVDZ-JKFIT (nbf 768, naux 3712,
nocc 160), whose packed (P|mn) (8.8 GB) passes through the device in three
chunks of at most 4 GiB. The FP64-only and integer-only contexts build
bitwise the same J and K as the default context in their modes, refuse the
other modes, and the integer build with two column groups (the half-transform
in FP64 rows, then split) stays within its bound of FP64."""
import sys

import numpy as np
import pytest

from nopt_amd import _lib as L
from nopt_amd import geometry as G
from nopt_amd.jk import JKContext
from nopt_amd.molecule import Molecule

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def w32():
    sym, xyz = G.water_cluster(32, 1)
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    say("(H2O)32: nbf", m.nbf, "naux", m.naux, "nocc", m.nocc)
    eri, V = m.eri3_packed(), m.metric()
    say("  packed (P|mn)", eri.nbytes / 1e9, "GB")
    rng = np.random.default_rng(3)
    C, _ = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))
    Co = np.ascontiguousarray(C[:, :m.nocc].T) * 0.5
    D = Co.T @ Co
    with JKContext(m.nbf, m.naux) as ctx:
        ctx.set_B(eri, V)
        ref = {p: ctx.build(D, Co, m.nocc, prec=p) for p in ("fp64", "int8")}
    yield m, eri, V, D, Co, ref


@pytest.mark.parametrize("keep,prec,other", [("fp64", "fp64", "int8"), ("int8", "int8", "fp64")])
def test_single_form_context_matches_default(w32, keep, prec, other):
    m, eri, V, D, Co, ref = w32
    with JKContext(m.nbf, m.naux, keep=keep) as ctx:
        ctx.set_B(eri, V)
        J, K = ctx.build(D, Co, m.nocc, prec=prec)
        dj, dk = np.abs(J - ref[prec][0]).max(), np.abs(K - ref[prec][1]).max()
        say(keep, "context,", prec, "build vs default context: max|dJ|", dj, "max|dK|", dk)
        assert dj == 0.0 and dk == 0.0
        with pytest.raises(L.NoptError):
            ctx.build(D, Co, m.nocc, prec=other)
        if keep == "int8":
            with pytest.raises(L.NoptError):
                ctx.get_B()


def test_integer_two_groups_within_bound(w32):
    m, eri, V, D, Co, ref = w32
    (J64, K64), (J, K) = ref["fp64"], ref["int8"]
    dE = np.sum(D * (2 * (J - J64) - (K - K64)))
    say("(H2O)32 int8 vs fp64: max|dJ|", np.abs(J - J64).max(), "max|dK|", np.abs(K - K64).max(), "dE", dE)
    assert np.abs(J - J64).max() < 1e-8 * max(1.0, np.abs(J64).max())
    assert np.abs(K - K64).max() < 1e-7 * max(1.0, np.abs(K64).max())
    assert abs(dE) < 1e-6
