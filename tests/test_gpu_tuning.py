"""The split-count autotuner: results after tuning equal the untuned build up
to the order of the FP64 partial sums, the tuned counts are kept by the
context, and out-of-range counts are refused."""
import sys

import numpy as np
import pytest

from nopt_amd import _lib as L
from nopt_amd import geometry as G
from nopt_amd.jk import DeviceArray, JKContext
from nopt_amd.molecule import Molecule

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_autotune_keeps_results(prec):
    sym, xyz = G.water_cluster(6, 3)
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    rng = np.random.default_rng(5)
    C, _ = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))
    Co = np.ascontiguousarray(C[:, :m.nocc].T) * 0.5
    D = Co.T @ Co
    with JKContext(m.nbf, m.naux) as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        J0, K0 = ctx.build(D, Co, m.nocc, prec=prec)
        t0 = ctx.tuning()
        dD, dC = DeviceArray.from_host(D), DeviceArray.from_host(Co)
        t, ms = ctx.autotune(dD.ptr, dC.ptr, m.nocc, prec=prec, reps=2)
        say(prec, "nbf", m.nbf, "default", t0, "tuned", t, "ms/build", ms)
        assert ctx.tuning() == t and ms > 0
        J, K = ctx.build(D, Co, m.nocc, prec=prec)
        dj = np.abs(J - J0).max() / np.abs(J0).max()
        dk = np.abs(K - K0).max() / np.abs(K0).max()
        say("  after tuning: rel max|dJ|", dj, "rel max|dK|", dk)
        assert dj < 1e-13 and dk < 1e-13
        ctx.set_tuning(**t0)
        assert ctx.tuning() == t0
        with pytest.raises(L.NoptError):
            ctx.set_tuning(nsk8=0)
        with pytest.raises(L.NoptError):
            ctx.set_tuning(nsk=0)
        dD.free()
        dC.free()


def test_exchange_split_counts_not_multiples_of_8():
    """The integer exchange takes any Q-split count (the XCD ranges of its
    blocks then differ in size by one): K for 1, 3, 7 and 13 splits equals
    the 8-split K to the rounding of the FP64 sums over the splits, and a
    build repeated at one count is the same to the bit."""
    sym, xyz = G.water_cluster(6, 3)
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    rng = np.random.default_rng(7)
    C, _ = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))
    Co = np.ascontiguousarray(C[:, :m.nocc].T)
    D = Co.T @ Co
    with JKContext(m.nbf, m.naux) as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        t = ctx.tuning()
        ctx.set_tuning(**dict(t, nsk8=8))
        J8, K8 = ctx.build(D, Co, m.nocc, prec="int8")
        for ns in (1, 3, 7, 13):
            ctx.set_tuning(**dict(t, nsk8=ns))
            assert ctx.tuning()["nsk8"] == ns
            J, K = ctx.build(D, Co, m.nocc, prec="int8")
            J2, K2 = ctx.build(D, Co, m.nocc, prec="int8")
            dk = float(np.abs(K - K8).max() / np.abs(K8).max())
            say("nsk8", ns, "rel max|dK| vs 8 splits", dk)
            assert np.array_equal(K, K2) and np.array_equal(J, J2)
            assert np.array_equal(J, J8)
            assert dk < 1e-13
