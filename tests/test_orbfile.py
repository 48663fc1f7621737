"""NOPT's binary orbital files (.orb, src/binary_files.cpp:248-379): the
reader and writer, and NOPT's own RHF orbitals of p-benzoquinone
(examples/dmrg/pBQ/pBQ_2B2_RHF.orb, written by NOPT's RI-RHF of
pBQ_2A_RHF.inp in GAMESS order and normalisation) and of octatetraene
(examples/dmrg/4ene/4ene_RHF.orb) carried into this
package's Cartesian AO order: orthonormal in this package's overlap, and
their orbital energies, diagonal of the Fock matrix the oracle builds from
them, are the energies NOPT stored in the file."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from nopt_amd.molecule import Molecule
from nopt_amd.orbfile import MAGIC, gamess_to_libint, libint_to_gamess, read_orb, write_orb
from systems import PBQ, PBQ_ORB

BASIS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nopt_amd", "basis")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    C, e = rng.standard_normal((7, 9)), rng.standard_normal(7)
    p = tmp_path / "a.orb"
    write_orb(p, C, e)
    r = read_orb(p)
    assert (r.n_mo, r.n_ao) == (7, 9) and np.array_equal(r.C, C) and np.array_equal(r.eps, e)
    assert r.C_beta is None and os.path.getsize(p) == 4 * 3 + 8 * (63 + 7) + 4
    Cb, eb = rng.standard_normal((7, 9)), rng.standard_normal(7)
    write_orb(p, C, e, Cb, eb)
    r = read_orb(p)
    assert np.array_equal(r.C_beta, Cb) and np.array_equal(r.eps_beta, eb)


def test_reader_rejects_bad_files(tmp_path):
    p = tmp_path / "b.orb"
    write_orb(p, np.ones((2, 3)), np.ones(2))
    raw = bytearray(open(p, "rb").read())
    bad = tmp_path / "bad.orb"
    bad.write_bytes(bytes([0, 0, 0, 0]) + bytes(raw[4:]))  # wrong mark
    with pytest.raises(ValueError):
        read_orb(bad)
    bad.write_bytes(bytes(raw[:-9]))  # truncated
    with pytest.raises(ValueError):
        read_orb(bad)
    assert int(np.frombuffer(bytes(raw[:4]), dtype=np.int32)[0]) == MAGIC


@pytest.fixture(scope="module")
def pbq():
    m = Molecule.from_angstrom(*PBQ, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    f = read_orb(PBQ_ORB)
    return m, f


def test_gamess_and_libint_forms_are_inverse(pbq):
    m, f = pbq
    C = gamess_to_libint(f.C, m.function_l())
    assert np.allclose(libint_to_gamess(C, m.function_l()), f.C, rtol=0, atol=1e-15)


def test_nopt_orbitals_are_orthonormal_in_our_basis(pbq):
    m, f = pbq
    assert (f.n_mo, f.n_ao, m.nbf) == (132, 140, 140) and f.C_beta is None
    S = m.one_electron()[0]
    C = gamess_to_libint(f.C, m.function_l())
    dev = np.abs(C @ S @ C.T - np.eye(f.n_mo)).max()
    raw = np.abs(f.C @ S @ f.C.T - np.eye(f.n_mo)).max()  # without the reordering
    say("max|C S C^T - I|", dev, "as stored (GAMESS form):", raw)
    assert dev < 1e-10
    assert raw > 1e-2


def test_nopt_orbital_energies_from_oracle_fock(pbq):
    """F = H + 2J - K from the oracle's NOPT-path J/K at the file's occupied
    orbitals: diag(C F C^T) equals the file's orbital energies (NOPT converged
    to grad 1e-8, pBQ_2A_RHF.inp:20), and the energy of these orbitals equals
    the oracle's own converged RI-RHF energy."""
    m, f = pbq
    om = O.OracleMolecule.from_angstrom(PBQ[0], PBQ[1], os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-jkfit.gbs"))
    S, T, V = om.one_electron()
    H = T + V
    C = gamess_to_libint(f.C, m.function_l())
    Co = C[:om.nocc]
    D = Co.T @ Co
    jk = O.NoptJK(om.eri3_nopt(), O.s05(om.metric())[0], om.nocc)
    J, K = jk(D, Co, om.nocc)
    F = H + 2 * J - K
    eps = np.einsum("im,mn,in->i", C, F, C)
    E = float(np.sum(D * F) + np.sum(D * H) + om.e_nuc)
    E_scf = O.rhf(om)["E"]
    say("max|eps - eps_file|", np.abs(eps - f.eps).max(), "E(file orbitals)", E, "oracle RHF", E_scf)
    assert np.abs(eps - f.eps).max() < 2e-6
    assert abs(E - E_scf) < 1e-9


def test_nopt_octatetraene_orbitals():
    """NOPT's RHF orbitals of octatetraene (examples/dmrg/4ene/4ene_RHF.orb,
    from 4ene_RHF.inp: cc-pVDZ / cc-pVDZ-JKFIT, D5=1, 162 spherical orbitals
    over 170 Cartesian AOs): orthonormal after the reordering, their energies
    are the diagonal of the oracle's NOPT-path Fock matrix at them (NOPT
    converged to grad 1e-8, 4ene_RHF.inp:29), and its occupied-virtual block
    vanishes to the same order."""
    from systems import ENE4, ENE4_ORB
    m = Molecule.from_angstrom(*ENE4, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=False)
    f = read_orb(ENE4_ORB)
    assert (f.n_mo, f.n_ao, m.nbf, m.nocc) == (162, 170, 170, 29) and f.C_beta is None
    C = gamess_to_libint(f.C, m.function_l())
    om = O.OracleMolecule.from_angstrom(ENE4[0], ENE4[1], os.path.join(BASIS, "cc-pvdz.gbs"),
                                        os.path.join(BASIS, "cc-pvdz-jkfit.gbs"))
    S, T, V = om.one_electron()
    dev = np.abs(C @ S @ C.T - np.eye(f.n_mo)).max()
    H = T + V
    Co = C[:om.nocc]
    D = Co.T @ Co
    J, K = O.NoptJK(om.eri3_nopt(), O.s05(om.metric())[0], om.nocc)(D, Co, om.nocc)
    F = H + 2 * J - K
    eps = np.einsum("im,mn,in->i", C, F, C)
    Fai = np.abs(C[:om.nocc] @ F @ C[om.nocc:].T).max()
    say("max|C S C^T - I|", dev, "max|eps - eps_file|", np.abs(eps - f.eps).max(), "max|F_ai|", Fai)
    assert dev < 1e-10
    assert np.abs(eps - f.eps).max() < 5e-6
    assert Fai < 1e-6
