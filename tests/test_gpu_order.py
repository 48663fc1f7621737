"""Ordering and lifetime rules of a context's work buffers.

All builds of a context share its work buffers, so calls on different
streams must run one after the other, and a captured HIP graph that holds the
buffers' addresses must keep them from moving (include/nopt_jk.h)."""
import sys

import numpy as np
import pytest

from nopt_amd import _lib as L
from nopt_amd.molecule import Molecule
from nopt_amd.jk import JKContext, DeviceArray, Graph, synchronize
from systems import SYSTEMS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _orbitals(nbf, nocc, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    C, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    Co = np.ascontiguousarray(C[:, :nocc].T) * scale
    return Co, Co.T @ Co


@pytest.fixture(scope="module")
def guanine():
    sym, xyz = SYSTEMS["guanine"]
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    ctx = JKContext(m.nbf, m.naux)
    ctx.set_B(m.eri3_packed(), m.metric())
    yield m, ctx
    ctx.close()


@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_device_build_then_host_build_without_sync(guanine, prec):
    """A device build on the default stream followed at once by a host build
    (the context's own stream) of another density: both equal the builds run
    one at a time with a synchronisation between them."""
    m, ctx = guanine
    n = m.nbf
    C1, D1 = _orbitals(n, m.nocc, 1)
    C2, D2 = _orbitals(n, m.nocc, 2, scale=0.2)
    J1r, K1r = ctx.build(D1, C1, m.nocc, prec=prec)
    J2r, K2r = ctx.build(D2, C2, m.nocc, prec=prec)
    dD, dC = DeviceArray.from_host(D1), DeviceArray.from_host(C1)
    dJ, dK = DeviceArray((n, n)), DeviceArray((n, n))
    for rep in range(3):
        ctx.build_device(dD.ptr, dC.ptr, m.nocc, dJ.ptr, dK.ptr, prec, stream=0)
        J2, K2 = ctx.build(D2, C2, m.nocc, prec=prec)  # no synchronisation in between
        synchronize()
        J1, K1 = dJ.to_host(), dK.to_host()
        d = [float(np.abs(a - b).max()) for a, b in ((J1, J1r), (K1, K1r), (J2, J2r), (K2, K2r))]
        say(prec, "rep", rep, "max|dJ1| max|dK1| max|dJ2| max|dK2|", d)
        assert max(d) == 0.0  # same kernels, same order of sums: bitwise


def test_graph_pins_its_contexts():
    """After capture, a build that would grow a captured context's buffers,
    a tuning change and autotuning are refused (NOPT_ESTATE) before any device
    work; destroying the context is deferred until the graph is destroyed, and
    the graph still replays correctly in between."""
    sym, xyz = SYSTEMS["C2H4"]
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-ri")
    n = m.nbf
    ctx = JKContext(n, m.naux)
    ctx.set_B(m.eri3_packed(), m.metric())
    C, D = _orbitals(n, m.nocc, 3)
    Cbig, _ = _orbitals(n, 40, 4)
    dD, dC = DeviceArray.from_host(D), DeviceArray.from_host(C)
    dCb = DeviceArray.from_host(Cbig)
    dJ, dK = DeviceArray((n, n)), DeviceArray((n, n))
    s = ctx.stream()
    ctx.build_device(dD.ptr, dC.ptr, m.nocc, dJ.ptr, dK.ptr, "fp64", stream=s)
    synchronize()
    J0, K0 = dJ.to_host(), dK.to_host()
    with Graph.capture(s) as cap:
        ctx.build_device(dD.ptr, dC.ptr, m.nocc, dJ.ptr, dK.ptr, "fp64", stream=s)
    g = cap.graph
    # 40 occupied orbitals need a wider half-transformed tensor than 8
    with pytest.raises(L.NoptError) as e:
        ctx.build_device(dD.ptr, dCb.ptr, 40, dJ.ptr, dK.ptr, "fp64", stream=s)
    say("grow:", e.value)
    assert e.value.code == L.NOPT_ESTATE
    with pytest.raises(L.NoptError) as e:
        ctx.set_tuning(nsk=max(1, ctx.tuning()["nsk"] // 2))
    assert e.value.code == L.NOPT_ESTATE
    with pytest.raises(L.NoptError) as e:
        ctx.autotune(dD.ptr, dC.ptr, m.nocc, "fp64", reps=1, stream=s)
    assert e.value.code == L.NOPT_ESTATE
    # builds of the captured shape still run eagerly
    ctx.build_device(dD.ptr, dC.ptr, m.nocc, dJ.ptr, dK.ptr, "fp64", stream=s)
    synchronize()
    assert np.abs(dJ.to_host() - J0).max() == 0.0
    ctx.close()  # deferred: the graph still refers to the buffers
    for rep in range(2):
        dJ.copy_from(np.zeros((n, n)))
        dK.copy_from(np.zeros((n, n)))
        g.launch(s)
        synchronize()
        dj, dk = np.abs(dJ.to_host() - J0).max(), np.abs(dK.to_host() - K0).max()
        say("replay after destroy", rep, dj, dk)
        assert dj == 0.0 and dk == 0.0
    g.close()  # frees the context as well
