"""Host logic of the SCF drivers on the CPU: rhf() and minimize_orbitals()
driven by a numpy stand-in for the GPU context (the same J/K/F/E/adjoint
contract as nopt_amd.jk.JKContext), on H2O/cc-pVDZ/cc-pVDZ-JKFIT."""
import numpy as np
import pytest
import scipy.linalg as sl

from nopt_amd.molecule import Molecule, unpack_sym
from nopt_amd import geometry as G
from nopt_amd.scf import rhf, minimize_orbitals, rotate


class NumpyJK:
    """J, K, Fock/energy and adjoint from the fitted B = L^-1 (P|mn), in numpy."""

    def __init__(self, m):
        Bp = sl.solve_triangular(np.linalg.cholesky(m.metric()), m.eri3_packed(), lower=True)
        self.B = unpack_sym(Bp, m.nbf)

    def build(self, D, C, nocc, prec="fp64"):
        J = np.einsum("qmn,q->mn", self.B, np.einsum("qmn,mn->q", self.B, D))
        X = np.einsum("qmn,in->qmi", self.B, C[:nocc])
        return J, np.einsum("qmi,qni->mn", X, X)

    def fock_energy(self, H, D, J, K, cj=2.0, ck=1.0):
        F = H + cj * J - ck * K
        return F, float(np.sum(D * (H + F)))

    def energy_gradient(self, H, C, nocc=None, prec="fp64", fock=False):
        nocc = nocc or C.shape[0]
        D = C[:nocc].T @ C[:nocc]
        F, E = self.fock_energy(H, D, *self.build(D, C, nocc))
        G_ = 4.0 * C[:nocc] @ F
        return (E, G_, F) if fock else (E, G_)


@pytest.fixture(scope="module")
def water():
    m = Molecule.from_angstrom(*G.water(), basis="cc-pvdz", aux="cc-pvdz-jkfit")
    S, T, V = m.one_electron()
    return m, NumpyJK(m), S, T + V


def core_guess(S, H):
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    return X @ np.linalg.eigh(X.T @ H @ X)[1]


def test_rhf_converges_and_is_stationary(water):
    m, ctx, S, H = water
    res = rhf(m, ctx=ctx, H=H, S=S, e_conv=1e-12, d_conv=1e-10)
    assert res.converged
    Co = res.C[:, :m.nocc]
    D = Co @ Co.T
    F, _ = ctx.fock_energy(H, D, *ctx.build(D, Co.T, m.nocc))
    assert np.abs(F @ D @ S - S @ D @ F).max() < 1e-7  # Brillouin: [F, D] = 0 in the S metric
    assert np.abs(res.C.T @ S @ res.C - np.eye(m.nbf)).max() < 1e-10


def test_direct_minimisation_matches_rhf(water):
    m, ctx, S, H = water
    ref = rhf(m, ctx=ctx, H=H, S=S, e_conv=1e-12, d_conv=1e-10)
    res = minimize_orbitals(m, ctx, H, core_guess(S, H), g_conv=1e-7, e_conv=1e-12)
    assert res.converged and abs(res.E - ref.E) < 1e-10
    # energies along the accepted path never rise
    Es = [h[1] for h in res.history]
    assert all(b <= a + 1e-12 for a, b in zip(Es, Es[1:]))


def test_rotation_keeps_orthonormality_and_gradient_is_directional_derivative(water):
    m, ctx, S, H = water
    C = core_guess(S, H)
    rng = np.random.default_rng(3)
    k = 0.1 * rng.standard_normal((m.nbf - m.nocc, m.nocc))
    Cr = rotate(C, m.nocc, k)
    assert np.abs(Cr.T @ S @ Cr - np.eye(m.nbf)).max() < 1e-12
    # dE/dt of E(C exp(t kappa)) at t=0 equals sum g_ai k_ai with g = 4 F_ai
    E0, Gr, F = ctx.energy_gradient(H, np.ascontiguousarray(C[:, :m.nocc].T), m.nocc, fock=True)
    g = 4.0 * (C.T @ F @ C)[m.nocc:, :m.nocc]
    h = 1e-5
    ep = ctx.energy_gradient(H, np.ascontiguousarray(rotate(C, m.nocc, h * k)[:, :m.nocc].T), m.nocc)[0]
    em = ctx.energy_gradient(H, np.ascontiguousarray(rotate(C, m.nocc, -h * k)[:, :m.nocc].T), m.nocc)[0]
    assert abs((ep - em) / (2 * h) - np.sum(g * k)) < 1e-6 * max(1.0, abs(np.sum(g * k)))
