"""Blocks of exact zeros in B's integer forms and the order of the functions.

The integer forms store an element below its row's resolution as an exact
zero; the half-transform leaves out the 64-column slabs, and the Coulomb pair
the segments, that hold nothing else, and the forms keep the functions in a
band order computed from the three-index tensor, so that such blocks gather.
None of this may change K: its products are exact integer sums, whatever
order the functions come in and whichever zero blocks are skipped. Relabelling
the basis functions of the input must therefore give the relabelled K to the
bit, and J to the rounding of its FP64 sums (gamma and J add their terms in
the functions' order)."""
import sys

import numpy as np
import pytest

from nopt_amd import geometry as G
from nopt_amd.jk import JKContext
from nopt_amd.molecule import Molecule

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relabel_packed(T, q):
    """T[P][pair(m, n)] of the functions relabelled so that new function a is
    old function q[a]."""
    n = len(q)
    r, c = np.tril_indices(n)
    a, b = q[r], q[c]
    hi, lo = np.maximum(a, b), np.minimum(a, b)
    return np.ascontiguousarray(T[:, hi * (hi + 1) // 2 + lo])


def test_relabelled_basis_gives_the_same_k():
    sym, xyz = G.alkane(8)
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    n = m.nbf
    eri3, V = m.eri3_packed(), m.metric()
    rng = np.random.default_rng(0)
    q = rng.permutation(n)
    nocc = m.nocc
    C = np.linalg.qr(rng.standard_normal((n, nocc)))[0].T.copy()
    D = C.T @ C
    out = {}
    for name, T, Cx, Dx in (("input order", eri3, C, D),
                            ("relabelled", relabel_packed(eri3, q), np.ascontiguousarray(C[:, q]),
                             np.ascontiguousarray(D[np.ix_(q, q)]))):
        with JKContext(n, m.naux) as ctx:
            ctx.set_B(T, V)
            shares = ctx.zero_blocks()
            J8, K8 = ctx.build(Dx, Cx, nocc, prec="int8")
            J64, K64 = ctx.build(Dx, Cx, nocc, prec="fp64")
        say(name, "shares read", shares, "max|dK| vs FP64", float(np.abs(K8 - K64).max()))
        assert shares["slabs"] < 1.0 and shares["segments"] < 1.0
        assert np.abs(K8 - K64).max() < 1e-7 * max(1.0, np.abs(K64).max())
        out[name] = (J8, K8)
    (J, K), (Jq, Kq) = out["input order"], out["relabelled"]
    dJ = float(np.abs(Jq - J[np.ix_(q, q)]).max())
    say("relabelled vs input order: K equal", np.array_equal(Kq, K[np.ix_(q, q)]), "max|dJ|", dJ)
    assert np.array_equal(Kq, K[np.ix_(q, q)])
    # J: gamma_Q sums some 10^4 products in FP64 and J adds 10^3 functions'
    # gamma_Q B_Q, in an order that follows the functions' order; measured
    # 3e-11 relative for this case
    assert dJ < 1e-9 * np.abs(J).max()


def test_fp64_only_context_reads_every_block():
    sym, xyz = G.water()
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    with JKContext(m.nbf, m.naux, keep="fp64") as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        assert ctx.zero_blocks() == dict(segments=1.0, slots=1.0, slabs=1.0)
