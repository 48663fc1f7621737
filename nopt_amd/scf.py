"""Closed-shell RI-RHF driver around the GPU RI-J/K build.

The loop mirrors NOPT's RHF (src/SCF.cpp:78-223): per iteration the density
D = C_occ C_occ^T (gen_HF_DM, src/SCF.cpp:63-73), one RI-J/K build and
F = H + 2J - K (calc_F_AO, src/molecule.cpp:4407-4438), the energy
E = sum D.F + sum D.H + V_nuc (src/SCF.cpp:134) and the orbital gradient
4 F_ai (src/converger_2_1.cpp:54-62). NOPT steps with SOSCF/L-BFGS from a
Hueckel or SAD guess; this driver uses a core-Hamiltonian guess and DIIS on
FC S - S C F... (the commutator FDS - SDF), which converges to the same fixed
point. Orthogonalisation (S^-1/2), diagonalisation and the energy stay FP64.

Mixed precision as iterative refinement: builds use `prec` (integer split,
FP32 or bf16 exchange) until the DIIS error falls below `switch_err`, then
FP64 for the remaining iterations, so the converged energy is that of the
FP64 build.
"""
import time

import numpy as np

from .jk import JKContext


class SCFResult(dict):
    __getattr__ = dict.__getitem__


def rhf(mol, ctx=None, prec="fp64", switch_err=1e-5, e_conv=1e-10, d_conv=1e-8, max_it=100,
        diis_size=8, verbose=False, H=None, S=None, switch_ctx=None):
    """Returns SCFResult(E, C (nbf x nbf, columns are MOs), eps, D, F, iters,
    builds {prec: count}, history, switch_iter). switch_ctx, if given, is
    called with the current context when the FP64 end-game starts and returns
    the context for the remaining builds (a large system that holds only the
    integer form of B in HBM hands over to an FP64 context there)."""
    own = ctx is None
    if own:
        ctx = JKContext(mol.nbf, mol.naux)
        ctx.set_B(mol.eri3_packed(), mol.metric())
    try:
        if H is None or S is None:
            S_, T_, V_ = mol.one_electron()
            S = S_ if S is None else S
            H = (T_ + V_) if H is None else H
        s, U = np.linalg.eigh(S)
        X = U @ np.diag(s ** -0.5) @ U.T
        eps, Cp = np.linalg.eigh(X.T @ H @ X)
        C = X @ Cp
        nocc = mol.nocc
        cur = prec
        E_old, D_old = 0.0, None
        fs, es = [], []
        builds, hist = {}, []
        switch_iter = None
        for it in range(max_it):
            Co = C[:, :nocc]
            D = Co @ Co.T
            J, K = ctx.build(D, Co.T, nocc, prec=cur)
            built = cur  # precision of the build that gave this iteration's energy
            builds[cur] = builds.get(cur, 0) + 1
            F, e_el = ctx.fock_energy(H, D, J, K, 2.0, 1.0)
            E = e_el + mol.e_nuc
            err = F @ D @ S - S @ D @ F
            emax = float(np.abs(err).max())
            drms = float(np.sqrt(np.mean((D - D_old) ** 2))) if D_old is not None else np.inf
            hist.append((it, cur, E, emax, drms))
            if verbose:
                print(f"ITER {it} {cur:7s} E={E:.12f} dE={E - E_old:.3e} err={emax:.3e} drms={drms:.3e}")
            done = abs(E - E_old) < e_conv and drms < d_conv
            if done and cur == "fp64":
                break
            if cur != "fp64" and (emax < switch_err or done):
                cur = "fp64"  # FP64 end-game: refine with exact exchange
                fs, es = [], []
                switch_iter = it
                if switch_ctx is not None:
                    ctx = switch_ctx(ctx)
            E_old, D_old = E, D
            fs.append(F)
            es.append(err)
            if len(fs) > diis_size:
                fs.pop(0)
                es.pop(0)
            n = len(fs)
            Bm = -np.ones((n + 1, n + 1))
            Bm[-1, -1] = 0.0
            for i in range(n):
                for j in range(i, n):
                    Bm[i, j] = Bm[j, i] = np.sum(es[i] * es[j])
            rhs = np.zeros(n + 1)
            rhs[-1] = -1.0
            cf = np.linalg.lstsq(Bm, rhs, rcond=None)[0][:n]
            Fd = sum(c * f for c, f in zip(cf, fs))
            eps, Cp = np.linalg.eigh(X.T @ Fd @ X)
            C = X @ Cp
        # converged only on an FP64 iteration: a run that switches to the FP64
        # end-game at its last iteration has not refined its energy
        return SCFResult(E=E, C=C, eps=eps, D=D, F=F, H=H, S=S, iters=it, builds=builds, history=hist,
                         converged=bool(done and built == "fp64"), switch_iter=switch_iter)
    finally:
        if own:
            ctx.close()


def rotate(C, nocc, k):
    """C exp(kappa) with kappa the occupied-virtual antisymmetric generator
    whose virtual-occupied block is k (nvirt x nocc)."""
    import scipy.linalg as sl
    n = C.shape[1]
    kap = np.zeros((n, n))
    kap[nocc:, :nocc] = k
    kap[:nocc, nocc:] = -k.T
    return C @ sl.expm(kap)


def minimize_orbitals(mol, ctx, H, C0, prec="fp64", g_conv=1e-6, e_conv=1e-11, max_it=300, memory=8,
                      trust=0.5, verbose=False):
    """Direct minimisation of E_RHF over orbital rotations C <- C exp(kappa)
    with the GPU energy/adjoint (dE/dC_occ = 4 F C_occ), in the way of NOPT's
    SOSCF/L-BFGS engine (src/converger_2_1.cpp:22-123, src/l-bfgs_2_1.cpp:33-125):
    gradient g_ai = 4 F_ai in the current orbitals, the diagonal Hessian
    estimate 4 (F_aa - F_ii) as preconditioner, L-BFGS directions within a
    trust radius on max|kappa|, the orbitals re-based after every step, and a
    step that raises the energy rejected (radius halved, history cleared).
    C0 must be orthonormal in the metric S. Returns SCFResult(E, C, iters,
    gmax, history, converged)."""
    nocc = mol.nocc
    C = np.array(C0, dtype=np.float64)

    def evaluate(Cc):
        Co = np.ascontiguousarray(Cc[:, :nocc].T)
        E_el, _, F = ctx.energy_gradient(H, Co, nocc, prec=prec, fock=True)
        Fmo = Cc.T @ F @ Cc
        return E_el + mol.e_nuc, 4.0 * Fmo[nocc:, :nocc], np.diag(Fmo)

    E, g, eps = evaluate(C)
    S_hist, Y_hist, hist = [], [], []
    E_prev = None
    for it in range(max_it):
        gmax = float(np.abs(g).max())
        hist.append((it, E, gmax, trust))
        if verbose:
            print(f"MIN {it} E={E:.12f} |g|max={gmax:.3e} trust={trust:.3f}")
        if gmax < g_conv and E_prev is not None and abs(E - E_prev) < e_conv:
            return SCFResult(E=E, C=C, iters=it, gmax=gmax, history=hist, converged=True)
        hdiag = np.maximum(4.0 * (eps[nocc:, None] - eps[None, :nocc]), 0.2)
        # L-BFGS two-loop recursion, H0 = diag(1/hdiag)
        q = g.copy()
        saved = []
        for s_, y_ in zip(reversed(S_hist), reversed(Y_hist)):
            rho = 1.0 / float(np.sum(y_ * s_))
            a = rho * float(np.sum(s_ * q))
            saved.append((a, rho, s_, y_))
            q -= a * y_
        r = q / hdiag
        for a, rho, s_, y_ in reversed(saved):
            r += (a - rho * float(np.sum(y_ * r))) * s_
        step = -r
        if float(np.sum(step * g)) >= 0.0:  # not a descent direction: preconditioned gradient
            step = -g / hdiag
            S_hist.clear()
            Y_hist.clear()
        big = float(np.abs(step).max())
        if big > trust:
            step *= trust / big
        Cn = rotate(C, nocc, step)
        En, gn, epsn = evaluate(Cn)
        if En > E + 1e-12 * max(1.0, abs(E)):
            trust = max(0.5 * trust, 1e-4)  # reject
            S_hist.clear()
            Y_hist.clear()
            continue
        y = gn - g
        if float(np.sum(step * y)) > 1e-14:
            S_hist.append(step)
            Y_hist.append(y)
            if len(S_hist) > memory:
                S_hist.pop(0)
                Y_hist.pop(0)
        if big <= trust:
            trust = min(2.0 * trust, 1.0)
        E_prev = E
        C, E, g, eps = Cn, En, gn, epsn
    return SCFResult(E=E, C=C, iters=max_it, gmax=float(np.abs(g).max()), history=hist, converged=False)
