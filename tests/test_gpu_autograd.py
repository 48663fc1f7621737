"""The torch autograd function over the GPU build: backward equals the
adjoint 4 C F, agrees with central differences, and a torch optimiser over
orbital rotations reaches the DIIS energy of H2O/cc-pVDZ/cc-pVDZ-JKFIT."""
import sys

import numpy as np
import pytest

from nopt_amd.molecule import Molecule
from nopt_amd.jk import JKContext
from nopt_amd import geometry as G
from nopt_amd.scf import rhf

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def water():
    from nopt_amd.autograd import rhf_energy  # noqa: F401  (import check)
    m = Molecule.from_angstrom(*G.water(), basis="cc-pvdz", aux="cc-pvdz-jkfit")
    S, T, V = m.one_electron()
    H = T + V
    jk = JKContext(m.nbf, m.naux)
    jk.set_B(m.eri3_packed(), m.metric())
    ref = rhf(m, ctx=jk, H=H, S=S, e_conv=1e-12, d_conv=1e-10)
    yield m, jk, S, H, ref
    jk.close()


def test_backward_is_the_adjoint(water):
    from nopt_amd.autograd import rhf_energy
    m, jk, S, H, ref = water
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0)
    Co = np.ascontiguousarray(ref.C[:, :m.nocc].T) + 1e-2 * rng.standard_normal((m.nocc, m.nbf))
    C = torch.tensor(Co, device=dev, requires_grad=True)
    Ht = torch.tensor(H, device=dev)
    E = rhf_energy(C, Ht, jk, e_nuc=m.e_nuc)
    E.backward()
    E_host, G_host = jk.energy_gradient(H, Co, m.nocc)
    say("E", float(E), "host", E_host + m.e_nuc, "max|grad - G|", float((C.grad.cpu() - torch.tensor(G_host)).abs().max()))
    assert abs(float(E) - (E_host + m.e_nuc)) < 1e-10
    assert float((C.grad.cpu() - torch.tensor(G_host)).abs().max()) < 1e-12
    d = torch.tensor(rng.standard_normal(Co.shape), device=dev)
    d /= d.norm()
    h = 1e-5
    with torch.no_grad():
        fd = (rhf_energy(C + h * d, Ht, jk) - rhf_energy(C - h * d, Ht, jk)) / (2 * h)
    an = float((C.grad * d).sum())
    say("directional: FD", float(fd), "analytic", an)
    assert abs(float(fd) - an) < 1e-7 * max(1.0, abs(an))


def test_torch_optimiser_reaches_diis_energy(water):
    from nopt_amd.autograd import rhf_energy, rotated_occupied
    m, jk, S, H, ref = water
    dev = torch.device("cuda:0")
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    C0 = torch.tensor(X @ np.linalg.eigh(X.T @ H @ X)[1], device=dev)
    Ht = torch.tensor(H, device=dev)
    kappa = torch.zeros((m.nbf - m.nocc, m.nocc), dtype=torch.float64, device=dev, requires_grad=True)
    opt = torch.optim.LBFGS([kappa], lr=1.0, max_iter=200, tolerance_grad=1e-9, tolerance_change=1e-14,
                            history_size=20, line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        E = rhf_energy(rotated_occupied(C0, kappa), Ht, jk, e_nuc=m.e_nuc)
        E.backward()
        return E

    for _ in range(5):
        E = float(opt.step(closure))
    say("torch L-BFGS E", E, "DIIS", ref.E, "diff", E - ref.E)
    assert abs(E - ref.E) < 1e-9
