"""GPU parity of the RI-J/K build against the oracle (FP64 restatement of
NOPT's DM_to_F_transform_RI, src/RI.cpp:115-176), through the C-ABI."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from nopt_amd.molecule import Molecule, unpack_sym
from nopt_amd.jk import JKContext
from nopt_amd import geometry as G
from systems import SYSTEMS, GOLDEN_E

pytestmark = pytest.mark.gpu
BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nopt_amd", "basis")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_case(sym, xyz_ang, aux):
    om = O.OracleMolecule.from_angstrom(sym, xyz_ang, os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, aux + ".gbs"))
    r = O.rhf(om)
    return om, r


@pytest.fixture(scope="module")
def h2o_case():
    sym, xyz = G.water()
    om, r = _oracle_case(sym, xyz, "cc-pvdz-jkfit")
    pm = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    return om, r, pm


@pytest.fixture(scope="module")
def c2h4_case():
    sym, xyz = SYSTEMS["C2H4"]
    om, r = _oracle_case(sym, xyz, "cc-pvdz-ri")
    pm = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-ri", pure=False)
    return om, r, pm


@pytest.mark.parametrize("case", ["h2o_case", "c2h4_case"])
def test_jk_fp64_matches_oracle(case, request):
    """Cartesian AOs, NOPT's own representation: element-wise J and K."""
    om, r, pm = request.getfixturevalue(case)
    say("case", case, "n_ao", pm.nbf, "naux", pm.naux)
    assert pm.nbf == om.n_ao and pm.naux == om.naux
    D, Cc = r["D"], r["C"]
    J0, K0 = r["jk"](D, Cc.T, om.nocc)
    with JKContext(pm.nbf, pm.naux) as ctx:
        ctx.set_B(pm.eri3_packed(), pm.metric())
        J, K = ctx.build(D, Cc.T, om.nocc, prec="fp64")
    say("max|dJ|", np.abs(J - J0).max(), "max|dK|", np.abs(K - K0).max())
    assert np.abs(J - J0).max() < 1e-10
    assert np.abs(K - K0).max() < 1e-10
    E0 = np.sum(D * (2 * r["H"] + 2 * J0 - K0))
    E1 = np.sum(D * (2 * r["H"] + 2 * J - K))
    assert abs(E1 - E0) < 1e-10


def test_B_formation_matches_host_solve(h2o_case):
    """Device Cholesky + triangular solve against scipy on the same integrals."""
    import scipy.linalg as sl
    _, _, pm = h2o_case
    A, V = pm.eri3_packed(), pm.metric()
    Bref = sl.solve_triangular(np.linalg.cholesky(V), A, lower=True)
    with JKContext(pm.nbf, pm.naux) as ctx:
        ctx.set_B(A, V)
        B = ctx.get_B()
    say("max|dB|", np.abs(B - Bref).max())
    assert np.abs(B - Bref).max() < 1e-11 * max(1.0, np.abs(Bref).max())


def test_metric_not_positive_definite_is_reported(h2o_case):
    from nopt_amd._lib import NoptError, NOPT_ECHOL
    _, _, pm = h2o_case
    V = pm.metric()
    V[3, 3] = -1.0  # indefinite: must be reported, not silently dropped
    with JKContext(pm.nbf, pm.naux) as ctx:
        with pytest.raises(NoptError) as ei:
            ctx.set_B(pm.eri3_packed(), V)
        assert ei.value.code == NOPT_ECHOL


@pytest.mark.parametrize("prec,tol", [("fp32", 5e-6), ("bf16x6", 5e-6), ("bf16x3", 2e-4), ("int8", 5e-8)])
def test_mixed_precision_close_to_fp64(c2h4_case, prec, tol):
    om, r, pm = c2h4_case
    D, Cc = r["D"], r["C"]
    with JKContext(pm.nbf, pm.naux) as ctx:
        ctx.set_B(pm.eri3_packed(), pm.metric())
        J64, K64 = ctx.build(D, Cc.T, om.nocc, prec="fp64")
        J, K = ctx.build(D, Cc.T, om.nocc, prec=prec)
    say(prec, "max|dJ|", np.abs(J - J64).max(), "max|dK|", np.abs(K - K64).max())
    assert np.abs(J - J64).max() < 1e-12 * max(1, np.abs(J64).max()) * 100
    assert np.abs(K - K64).max() < tol


@pytest.mark.parametrize("name", ["C2H4", "C6H8", "guanine"])
def test_spherical_rhf_reproduces_nopt_log_energy(name):
    """Full RI-RHF on the GPU build, spherical AOs: the energies NOPT prints
    for C2H4, C6H8 and guanine (cc-pVDZ / cc-pVDZ-RI)."""
    from nopt_amd.scf import rhf
    sym, xyz = SYSTEMS[name]
    m = Molecule.from_angstrom(sym, xyz, basis="cc-pvdz", aux="cc-pvdz-ri")
    res = rhf(m, e_conv=1e-11, d_conv=1e-9)
    say(name, "nbf", m.nbf, "naux", m.naux, "E", res.E, "ref", GOLDEN_E[name], "diff", res.E - GOLDEN_E[name])
    assert abs(res.E - GOLDEN_E[name]) < 3e-10


def test_energy_gradient_matches_finite_differences():
    """dE/dC = 4 C F against central differences of the GPU energy."""
    sym, xyz = G.water()
    m = Molecule.from_angstrom(sym, xyz)
    S, T, V = m.one_electron()
    H = T + V
    from nopt_amd.scf import rhf
    with JKContext(m.nbf, m.naux) as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        res = rhf(m, ctx=ctx)
        C = np.ascontiguousarray(res.C[:, :m.nocc].T)
        rng = np.random.default_rng(0)
        C = C + 1e-2 * rng.standard_normal(C.shape)  # off the stationary point
        E0, Gr = ctx.energy_gradient(H, C, prec="fp64")
        h = 1e-5
        fd = np.zeros_like(C)
        for (i, mu) in [(0, 0), (1, 3), (2, 7), (4, 11), (3, 20)]:
            Cp = C.copy(); Cp[i, mu] += h
            Cm = C.copy(); Cm[i, mu] -= h
            fd[i, mu] = (ctx.energy_gradient(H, Cp)[0] - ctx.energy_gradient(H, Cm)[0]) / (2 * h)
            say(i, mu, Gr[i, mu], fd[i, mu])
            assert abs(Gr[i, mu] - fd[i, mu]) < 1e-7 * max(1.0, abs(fd[i, mu]))


def test_nopt_ri_entry_points_match_oracle(c2h4_case):
    """NOPT's own call sequence through include/nopt_ri.h: hand over (P|Q) and
    RI_AA in NOPT's layout, then DM_to_F_transform_RI and the J/K half of
    calc_2el_MO_INTS_RI, against the oracle's restatement of src/RI.cpp:115-176."""
    from nopt_amd import _lib as L
    om, r, pm = c2h4_case
    lib = L.lib()
    D, Cc = r["D"], r["C"]
    MO = np.ascontiguousarray(Cc.T)
    J0, K0 = r["jk"](D, MO, om.nocc)
    V, A = pm.metric(), np.ascontiguousarray(pm.eri3_nopt())
    n = pm.nbf
    assert lib.RI_core_realloc(om.nocc, n) < 0          # no context yet
    assert lib.RI_gpu_attach(L.dptr(V), L.dptr(A), n, pm.naux, 0) == 0
    try:
        assert lib.RI_core_realloc(om.nocc, n) == 0
        J, K = np.zeros((n, n)), np.zeros((n, n))
        assert lib.DM_to_F_transform_RI(L.dptr(J), L.dptr(K), L.dptr(np.ascontiguousarray(D)), L.dptr(MO),
                                        om.nocc, n) == 1
        say("RI shim max|dJ|", np.abs(J - J0).max(), "max|dK|", np.abs(K - K0).max())
        assert np.abs(J - J0).max() < 1e-10 and np.abs(K - K0).max() < 1e-10
        J2, K2 = np.full((n, n), 7.0), np.full((n, n), 7.0)
        rc = lib.calc_2el_MO_INTS_RI(n, L.dptr(np.ascontiguousarray(D)), L.dptr(J2), L.dptr(K2), None,
                                     L.dptr(MO), None, None, om.nocc, 0, 0)
        assert rc == 0
        assert np.abs(J2 - J).max() == 0.0 and np.abs(K2 - K).max() == 0.0
        assert lib.calc_2el_MO_INTS_RI(n, None, None, None, None, None, None, None, om.nocc, 0, 0) == 0
        assert lib.DM_to_F_transform_RI(L.dptr(J), L.dptr(K), L.dptr(np.ascontiguousarray(D)), L.dptr(MO),
                                        om.nocc, n + 1) < 0
    finally:
        lib.regen_RI_AA_release()
    assert lib.DM_to_F_transform_RI(L.dptr(J), L.dptr(K), L.dptr(np.ascontiguousarray(D)), L.dptr(MO),
                                    om.nocc, n) < 0


def test_ri_h_cxx_names_build_like_the_c_entry_points(c2h4_case):
    """The C++-linkage functions with NOPT's include/RI.h:10-16 signatures
    (what NOPT's callers link to) give J, K and MO_INTS bitwise equal to the
    C entry points of nopt_ri.h, and J/K equal to the oracle's restatement."""
    import ctypes
    from nopt_amd import _lib as L
    om, r, pm = c2h4_case
    lib = L.lib()
    dp, lg = L._c_double_p, ctypes.c_long
    f_real = getattr(lib, "_Z15RI_core_reallocll")
    f_real.restype, f_real.argtypes = ctypes.c_int, [lg, lg]
    f_dm = getattr(lib, "_Z20DM_to_F_transform_RIPdS_S_S_ll")
    f_dm.restype, f_dm.argtypes = ctypes.c_int, [dp, dp, dp, dp, lg, lg]
    f_mo = getattr(lib, "_Z19calc_2el_MO_INTS_RIlPdS_S_S_S_S_S_lll")
    f_mo.restype, f_mo.argtypes = ctypes.c_int, [lg] + [dp] * 7 + [lg, lg, lg]
    D, Cc = r["D"], r["C"]
    MO = np.ascontiguousarray(Cc.T)
    D = np.ascontiguousarray(D)
    J0, K0 = r["jk"](D, MO, om.nocc)
    n, na = pm.nbf, 5
    V1 = np.ascontiguousarray(np.random.default_rng(9).standard_normal((na, n)))
    V, A = pm.metric(), np.ascontiguousarray(pm.eri3_nopt())
    assert lib.RI_gpu_attach(L.dptr(V), L.dptr(A), n, pm.naux, 0) == 0
    try:
        assert f_real(om.nocc, n) == 0
        Jc, Kc = np.zeros((n, n)), np.zeros((n, n))
        Jx, Kx = np.zeros((n, n)), np.zeros((n, n))
        assert lib.DM_to_F_transform_RI(L.dptr(Jc), L.dptr(Kc), L.dptr(D), L.dptr(MO), om.nocc, n) == 1
        assert f_dm(L.dptr(Jx), L.dptr(Kx), L.dptr(D), L.dptr(MO), om.nocc, n) == 1
        Ic, Ix = np.zeros((na,) * 4), np.zeros((na,) * 4)
        J2, K2 = np.zeros((n, n)), np.zeros((n, n))
        assert lib.calc_2el_MO_INTS_RI(n, None, None, None, L.dptr(Ic), None, L.dptr(V1), L.dptr(V1), 0, na, 0) == 0
        assert f_mo(n, L.dptr(D), L.dptr(J2), L.dptr(K2), L.dptr(Ix), L.dptr(MO), L.dptr(V1), L.dptr(V1),
                    om.nocc, na, 0) == 0
    finally:
        lib.regen_RI_AA_release()
    say("C++ names: max|dJ| vs oracle", np.abs(Jx - J0).max(), "max|dK|", np.abs(Kx - K0).max())
    assert np.array_equal(Jx, Jc) and np.array_equal(Kx, Kc)
    assert np.array_equal(J2, Jc) and np.array_equal(K2, Kc)
    assert np.array_equal(Ix, Ic)
    assert np.abs(Jx - J0).max() < 1e-10 and np.abs(Kx - K0).max() < 1e-10
    assert f_real(om.nocc, n) < 0  # released: no context


@pytest.mark.parametrize("na", [6, 9, 16, 17, 32])
@pytest.mark.parametrize("t", [0, 1])
def test_nopt_mo_ints_through_the_shim(c2h4_case, t, na):
    """The MO_INTS half of calc_2el_MO_INTS_RI through include/nopt_ri.h
    (orbitals as rows, t = 0, or as columns, t = 1; two different random
    active sets of na orbitals, covering each of the kernel's 8-, 16- and
    32-wide variants and both sides of their limits) against the oracle's
    restatement of src/RI.cpp:197-226, and nopt_jk_active_ints through
    JKContext against the same; J and K stay untouched when they are NULL."""
    from nopt_amd import _lib as L
    om, r, pm = c2h4_case
    lib = L.lib()
    n = pm.nbf
    rng = np.random.default_rng(5 + na)
    V1, V2 = rng.standard_normal((na, n)), rng.standard_normal((na, n))
    ref = O.nopt_mo_ints(om.eri3_nopt(), O.s05(om.metric())[0], V1, V2)
    V, A = pm.metric(), np.ascontiguousarray(pm.eri3_nopt())
    assert lib.RI_gpu_attach(L.dptr(V), L.dptr(A), n, pm.naux, 0) == 0
    try:
        out = np.full((na, na, na, na), 7.0)
        a1 = np.ascontiguousarray(V1.T if t else V1)
        a2 = np.ascontiguousarray(V2.T if t else V2)
        assert lib.calc_2el_MO_INTS_RI(n, None, None, None, L.dptr(out), None, L.dptr(a1), L.dptr(a2), 0, na,
                                       t) == 0
    finally:
        lib.regen_RI_AA_release()
    with JKContext(n, pm.naux) as ctx:
        ctx.set_B(pm.eri3_packed(), pm.metric())
        out2 = ctx.active_ints(V1, V2)
    scale = np.abs(ref).max()
    say("MO_INTS na", na, "t", t, "shim rel", np.abs(out - ref).max() / scale, "JKContext rel", np.abs(out2 - ref).max() / scale)
    assert np.abs(out - ref).max() < 1e-10 * scale
    assert np.abs(out2 - ref).max() < 1e-10 * scale


def test_active_ints_rejects_bad_requests(c2h4_case):
    """nopt_jk_active_ints: na outside 1..32 is NOPT_EINVAL and a context
    without the FP64 form of B is NOPT_ESTATE, each before any device work;
    the context stays usable afterwards."""
    from nopt_amd import _lib as L
    om, r, pm = c2h4_case
    n = pm.nbf
    C = np.random.default_rng(1).standard_normal((33, n))
    with JKContext(n, pm.naux) as ctx:
        ctx.set_B(pm.eri3_packed(), pm.metric())
        for bad in (C[:0], C[:33]):
            with pytest.raises(L.NoptError) as ei:
                ctx.active_ints(bad)
            assert ei.value.code == L.NOPT_EINVAL
        ok = ctx.active_ints(C[:2])
        assert ok.shape == (2, 2, 2, 2) and np.isfinite(ok).all()
    with JKContext(n, pm.naux, keep="int8") as ctx:
        ctx.set_B(pm.eri3_packed(), pm.metric())
        with pytest.raises(L.NoptError) as ei:
            ctx.active_ints(C[:2])
        assert ei.value.code == L.NOPT_ESTATE


# (system, orbital file, tolerance on the orbital energies): the file stores
# energies to NOPT's convergence, and the FP64 build at the file's orbitals
# differs from them by 5.3e-7 (pBQ) and 2.8e-6 (octatetraene, the oracle's
# NOPT-path build alike).
ORB_FILES = {"pBQ": ("PBQ", "PBQ_ORB", 2e-6), "4ene": ("ENE4", "ENE4_ORB", 5e-6)}


@pytest.mark.parametrize("name", sorted(ORB_FILES))
@pytest.mark.parametrize("prec", ["fp64", "int8"])
def test_nopt_orbital_file_through_the_gpu_build(name, prec):
    """NOPT's RHF orbitals of p-benzoquinone and octatetraene (their .orb
    files, GAMESS form, Cartesian AOs): the GPU build at those orbitals gives
    the orbital energies stored in the file, a vanishing occupied-virtual Fock
    block, and the energies of the oracle's NOPT-path build and of the same
    fit in Cholesky form."""
    import oracle as O
    import systems
    from nopt_amd.orbfile import read_orb, gamess_to_libint
    geo_name, orb_name, tol = ORB_FILES[name]
    geo, orb = getattr(systems, geo_name), getattr(systems, orb_name)
    m = Molecule.from_angstrom(*geo, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    f = read_orb(orb)
    C = gamess_to_libint(f.C, m.function_l())
    S, T, V = m.one_electron()
    H = T + V
    Co = np.ascontiguousarray(C[:m.nocc])
    D = Co.T @ Co
    with JKContext(m.nbf, m.naux) as ctx:
        ctx.set_B(m.eri3_packed(), m.metric())
        J, K = ctx.build(D, Co, m.nocc, prec=prec)
        E, _ = ctx.energy_gradient(H, Co, m.nocc, prec=prec)
    F = H + 2 * J - K
    eps = np.einsum("im,mn,in->i", C, F, C)
    # NOPT's orbitals are converged: the occupied-virtual block of the Fock
    # matrix in their basis vanishes to NOPT's threshold (3.2e-7 for pBQ and
    # 6.6e-7 for octatetraene with the oracle's NOPT-path build)
    Fai = np.abs((C[:m.nocc] @ F @ C[m.nocc:].T)).max()
    om = O.OracleMolecule.from_angstrom(geo[0], geo[1], os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-jkfit.gbs"))
    eri3, Vmet = om.eri3_nopt(), om.metric()
    Jo, Ko = O.NoptJK(eri3, O.s05(Vmet)[0], om.nocc)(D, Co, om.nocc)
    E_o = float(np.sum(D * (H + 2 * Jo - Ko)) + np.sum(D * H) + om.e_nuc)
    # the same fit in Cholesky form, B = L^-1 (P|mn) with V = L L^T, in FP64 on
    # the CPU: the form of the GPU build. NOPT's V^-1/2 form (eigendecomposition
    # of the JKFIT metric) amplifies rounding: its energy moved by 8.7e-10 Eh
    # between two hosts' BLAS, and sits 2.3e-10 (octatetraene) and 5.4e-10
    # (pBQ) from the Cholesky form; the GPU's FP64 energy equals the Cholesky
    # form's to all printed digits on the same host (4e-13 across hosts).
    n = m.nbf
    W = np.linalg.solve(np.linalg.cholesky(Vmet), eri3.reshape(n * n, -1).T).reshape(-1, n, n)
    Jc = np.tensordot(np.tensordot(W, D, axes=([1, 2], [0, 1])), W, axes=(0, 0))
    X = W @ Co.T
    Kc = np.einsum("pmi,pni->mn", X, X)
    E_c = float(np.sum(D * (H + 2 * Jc - Kc)) + np.sum(D * H) + om.e_nuc)
    say(name, prec, "max|eps - eps_file|", np.abs(eps - f.eps).max(), "max|F_ai|", Fai,
        "E", E + m.e_nuc, "oracle", E_o, "Cholesky form", E_c)
    assert np.abs(eps - f.eps).max() < tol
    assert Fai < max(1e-6, tol)
    assert abs(E + m.e_nuc - E_c) < (1e-11 if prec == "fp64" else 1e-6)
    assert abs(E + m.e_nuc - E_o) < (2e-9 if prec == "fp64" else 1e-6)
