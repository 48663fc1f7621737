"""The allocation-free build (nopt_jk_build_device_static, include/nopt_jk.h)
through the C-ABI: two calls in a row, the second with a smaller input, each
equal to nopt_jk_build_device's result; a build that needs more work space
than the static array holds is refused before any device work."""
import sys

import numpy as np
import pytest

from nopt_amd import _lib as L
from nopt_amd import geometry as G
from nopt_amd.jk import JKContext, DeviceArray, synchronize
from nopt_amd.molecule import Molecule
from systems import SYSTEMS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(sym, xyz, seed):
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit")
    ctx = JKContext(m.nbf, m.naux)
    ctx.set_B(m.eri3_packed(), m.metric())
    rng = np.random.default_rng(seed)
    C, _ = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))
    Co = np.ascontiguousarray(C[:, :m.nocc].T) * 0.4
    D = Co.T @ Co
    dev = dict(D=DeviceArray.from_host(D), C=DeviceArray.from_host(Co), J=DeviceArray((m.nbf, m.nbf)),
               K=DeviceArray((m.nbf, m.nbf)))
    return m, ctx, dev


@pytest.mark.parametrize("prec", ["int8", "fp64"])
def test_static_build_matches_build_device(prec):
    lib = L.lib()
    p = L.PREC[prec]
    say("static scratch bytes", lib.nopt_jk_static_scratch_bytes())
    big = _case(*SYSTEMS["guanine"], 1)  # nbf 171
    small = _case(*G.water(), 2)         # nbf 24: the second, smaller call
    for name, (m, ctx, d) in (("guanine", big), ("water", small)):
        say(prec, name, "nbf", m.nbf, "naux", m.naux, "nocc", m.nocc)
        L.check(lib.nopt_jk_build_device(ctx.handle, d["D"].ptr, d["C"].ptr, m.nocc, d["J"].ptr, d["K"].ptr, p,
                                         None))
        synchronize()
        J0, K0 = d["J"].to_host(), d["K"].to_host()
        d["J"].copy_from(np.zeros((m.nbf, m.nbf)))
        d["K"].copy_from(np.zeros((m.nbf, m.nbf)))
        L.check(lib.nopt_jk_build_device_static(ctx.handle, d["D"].ptr, d["C"].ptr, m.nocc, d["J"].ptr,
                                                d["K"].ptr, p, None))
        synchronize()
        J, K = d["J"].to_host(), d["K"].to_host()
        dj, dk = float(np.abs(J - J0).max()), float(np.abs(K - K0).max())
        say("  static vs build_device: max|dJ|", dj, "max|dK|", dk)
        assert dj == 0.0 and dk == 0.0  # same kernels and sums: bitwise
    # the larger system again after the smaller one (the array keeps the
    # last call's contents)
    m, ctx, d = big
    L.check(lib.nopt_jk_build_device_static(ctx.handle, d["D"].ptr, d["C"].ptr, m.nocc, d["J"].ptr, d["K"].ptr,
                                            p, None))
    synchronize()
    L.check(lib.nopt_jk_build_device(ctx.handle, d["D"].ptr, d["C"].ptr, m.nocc, d["J"].ptr, d["K"].ptr, p, None))
    synchronize()
    for _, c, dd in (big, small):
        for a in dd.values():
            a.free()
        c.close()


def test_static_build_refuses_oversized_work():
    """C20H42 in the FP64 mode needs a half-transformed tensor of at least
    8 naux nbf nocc bytes (0.75 GB), beyond the static array: NOPT_EINVAL, no
    device work; the context then builds normally."""
    lib = L.lib()
    m, ctx, d = _case(*G.alkane(20), 3)
    need = 8 * m.naux * m.nbf * m.nocc
    say("C20H42 FP64 X bytes at least", need, "scratch", lib.nopt_jk_static_scratch_bytes())
    assert need > lib.nopt_jk_static_scratch_bytes()
    rc = lib.nopt_jk_build_device_static(ctx.handle, d["D"].ptr, d["C"].ptr, m.nocc, d["J"].ptr, d["K"].ptr,
                                         L.PREC["fp64"], None)
    assert rc == L.NOPT_EINVAL
    say("refused:", lib.nopt_last_error().decode())
    L.check(lib.nopt_jk_build_device(ctx.handle, d["D"].ptr, d["C"].ptr, m.nocc, d["J"].ptr, d["K"].ptr,
                                     L.PREC["fp64"], None))
    synchronize()
    for a in d.values():
        a.free()
    ctx.close()
