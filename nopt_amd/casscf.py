"""State-averaged CASSCF around the GPU RI-J/K build: the caller of the
two-density J/K pair in NOPT's CASSCF (CAS_engine::calc_F, src/CAS.cpp:696-707;
nat_orb_calc, src/molecule.cpp:2925-2982; calc_F_AO, src/molecule.cpp:4407-4438;
the MO integrals of calc_2el_MO_INTS_RI, src/RI.cpp:178-226).

At orbitals C (columns: n_core core, n_act active, then virtual):
- the inactive Fock matrix F_I = H + 2 J(D_c) - K(D_c), D_c = C_core C_core^T
  (calc_F_AO with mc = 2: K from the n_core core rows);
- the active integrals (pu|vw) = sum_Q B_Q,pu B_Q,vw, p any orbital, u, v, w
  active, from the fitted B = L^-1 (P|mn) (NOPT's MO_INTS, src/RI.cpp:197-226,
  through its V^-1/2 fit);
- a determinant CI in the active space with the one-electron part
  C_act^T F_I C_act: the lowest `nstates` roots of spin S (mult = 2S + 1),
  their weighted (state-averaged) spin-summed 1- and 2-RDMs gamma and Gamma;
- the total Fock matrix F_I + F_A = H + J(2 D_c + D_a) - K([C_core; NO sqrt(n/2)]),
  D_a = C_act gamma C_act^T, NO and n the natural orbitals and occupations of
  gamma (CAS_engine::calc_F: d_o_DMV_gen, nat_orb_calc(g, 1, 'T'), then
  calc_F_AO with mc = 1, J from the density and K from the scaled rows);
- E = E_nuc + sum D_c (H + F_I) + sum_tu F_I,tu gamma_tu + 1/2 sum Gamma_tuvw (tu|vw),
  and the orbital gradient from the generalised Fock matrix
  G_iq = 2 (F_I + F_A)_qi (core i), G_tq = sum_u gamma_tu F_I,qu
  + sum_uvw Gamma_tuvw (qu|vw) (active t): dE/dkappa_pq = 2 (G_qp - G_pq)
  for C <- C exp(kappa), kappa antisymmetric (kappa_qp = -kappa_pq).

NOPT converges the orbitals with its SOSCF engine on the Lagrangian
(src/CAS.cpp:451-527); here preconditioned L-BFGS over the non-redundant
rotations (core-active, core-virtual, active-virtual; the active-active ones
leave a complete active space invariant) with the orbitals re-based after each
step. Both stop at the same stationary point of the averaged energy.

Every J/K goes through `jk(items) -> [(J, K), ...]`, items as for
JKContext.build_batch: (D, C rows, nocc). `gpu_jk(ctx, prec)` gives the GPU
build; the pair (D_c, C_core) and (2 D_c + D_a, [C_core; NO sqrt(n/2)]) is one
batched call when gamma is known (`pair_fock`).
"""
import itertools

import numpy as np
import scipy.linalg as sl
import scipy.sparse as sparse
import scipy.sparse.linalg as sla

from .molecule import unpack_sym


# ---------------------------------------------------------------- active CI
def _popcount(x):
    return bin(x).count("1")


def _excite(det, i, j):
    """a+_i a_j on a determinant (bit k = spin orbital k occupied, alpha
    spin orbitals 0..n-1 before beta n..2n-1): (sign, det') or None."""
    if not (det >> j) & 1:
        return None
    s = _popcount(det & ((1 << j) - 1))
    d = det ^ (1 << j)
    if (d >> i) & 1:
        return None
    s += _popcount(d & ((1 << i) - 1))
    return (-1.0 if s & 1 else 1.0), d | (1 << i)


class ActiveSpace:
    """Determinants of n_alpha + n_beta electrons in n orbitals and the
    spin-summed excitation operators E_pq (sparse) on CI vectors over them."""

    def __init__(self, n, n_alpha, n_beta):
        self.n, self.nel = n, n_alpha + n_beta
        strings = lambda k: [sum(1 << p for p in c) for c in itertools.combinations(range(n), k)]
        self.dets = [a | (b << n) for a in strings(n_alpha) for b in strings(n_beta)]
        self.index = {d: i for i, d in enumerate(self.dets)}
        self.ms = 0.5 * (n_alpha - n_beta)
        nd = len(self.dets)
        self.E = {}
        for p in range(n):
            for q in range(n):
                rows, cols, sg = [], [], []
                for k, d in enumerate(self.dets):
                    for off in (0, n):
                        r = _excite(d, p + off, q + off)
                        if r is not None:
                            rows.append(self.index[r[1]])
                            cols.append(k)
                            sg.append(r[0])
                self.E[p, q] = sparse.csr_matrix((sg, (rows, cols)), shape=(nd, nd))

    def apply(self, p, q, c):
        return self.E[p, q] @ c

    def spin_squared_matrix(self):
        """<I|S^2|J> as S_- S_+ + S_z^2 + S_z over the determinants (dense;
        small spaces)."""
        n, nd = self.n, len(self.dets)
        S2 = np.zeros((nd, nd))
        for k, d in enumerate(self.dets):
            up = {}
            for p in range(n):  # S_+ = sum_p a+_{p alpha} a_{p beta}
                r = _excite(d, p, p + n)
                if r is not None:
                    up[r[1]] = up.get(r[1], 0.0) + r[0]
            for du, cu in up.items():  # S_- = sum_q a+_{q beta} a_{q alpha}
                for q in range(n):
                    r = _excite(du, q + n, q)
                    if r is not None:
                        S2[self.index[r[1]], k] += cu * r[0]
            S2[k, k] += self.ms * self.ms + self.ms
        return S2

    def spin_squared(self, c):
        """<c|S^2|c> = N - N^2/4 - 1/2 sum_pq Gamma_pqqp for a normalised c,
        with sum_pq Gamma_pqqp = sum_pq |E_pq c|^2 - n N."""
        N, n = self.nel, self.n
        t = sum(float(np.dot(v, v)) for v in (self.E[p, q] @ c for p in range(n) for q in range(n)))
        return N - 0.25 * N * N - 0.5 * (t - n * N)

    def sigma(self, h, g, c):
        """(sum h_pq E_pq + 1/2 sum (pq|rs) (E_pq E_rs - d_qr E_ps)) c for a
        vector or the columns of a matrix; the d_qr term folds into the
        one-electron part as -1/2 sum_q (pq|qs)."""
        n = self.n
        hk = h - 0.5 * np.einsum("pqqs->ps", g)
        V = np.stack([self.E[r, s] @ c for r in range(n) for s in range(n)])  # row r*n+s: E_rs c
        out = np.tensordot(hk.reshape(n * n), V, axes=(0, 0))
        W = np.tensordot(0.5 * g.reshape(n * n, n * n), V, axes=(1, 0))
        for p in range(n):
            for q in range(n):
                out += self.E[p, q] @ W[p * n + q]
        return out

    def hamiltonian(self, h, g):
        """The CI matrix, dense (small spaces)."""
        Hm = self.sigma(h, g, np.eye(len(self.dets)))
        return 0.5 * (Hm + Hm.T)

    def rdms(self, c):
        """gamma_pq = <c|E_pq|c>, Gamma_pqrs = <c|E_pq E_rs|c> - d_qr gamma_ps
        (<c|E_pq E_rs|c> = <E_qp c|E_rs c>)."""
        n = self.n
        V = np.stack([self.E[r, s] @ c for r in range(n) for s in range(n)])  # row r*n+s
        gam = (V @ c).reshape(n, n)
        VT = V.reshape(n, n, -1).transpose(1, 0, 2).reshape(n * n, -1)    # row p*n+q = E_qp c
        Gam = (VT @ V.T).reshape(n, n, n, n)
        for q in range(n):
            Gam[:, q, q, :] -= gam
        return gam, Gam

    def solve(self, h, g, nstates, mult, weights=None, guess=None):
        """Lowest `nstates` roots of spin (mult - 1)/2: energies (without the
        core energy), vectors (columns) and the weighted RDMs. Dense
        diagonalisation up to 400 determinants, Lanczos (eigsh) beyond,
        asking for more roots until enough have the requested spin."""
        nd = len(self.dets)
        S = 0.5 * (mult - 1)
        # a spin the determinants cannot carry: S >= |Ms|, S at most half the
        # number of singly occupiable orbitals, 2S of the parity of the electrons
        s_max = 0.5 * min(self.nel, 2 * self.n - self.nel)
        if S < abs(self.ms) - 1e-9 or S > s_max + 1e-9 or (mult - 1 - self.nel) % 2:
            raise ValueError(f"no states of multiplicity {mult} for {self.nel} electrons in {self.n} orbitals "
                             f"with Ms = {self.ms}")
        if nd <= 400:
            w, V = np.linalg.eigh(self.hamiltonian(h, g))
        else:
            op = sla.LinearOperator((nd, nd), matvec=lambda x: self.sigma(h, g, x), dtype=np.float64)
            k = nstates + 4
            while True:
                # the last ground state, else a fixed-seed Gaussian vector (deterministic, and
                # with components along every state, which a uniform vector can lack)
                v0 = guess[:, 0] if guess is not None else np.random.default_rng(0).standard_normal(nd)
                w, V = sla.eigsh(op, k=k, which="SA", v0=v0, tol=1e-13)
                o = np.argsort(w)
                w, V = w[o], V[:, o]
                if sum(abs(self.spin_squared(V[:, j]) - S * (S + 1)) < 1e-6 for j in range(k)) >= nstates:
                    break
                if k >= nd - 1:  # every root eigsh can give: the spin filter left too few
                    break
                k = min(2 * k, nd - 1)
        pick = [j for j in range(len(w)) if abs(self.spin_squared(V[:, j]) - S * (S + 1)) < 1e-6][:nstates]
        if len(pick) < nstates:
            raise ValueError("fewer states of the requested spin than nstates")
        wts = np.ones(nstates) / nstates if weights is None else np.asarray(weights, float) / np.sum(weights)
        gam = np.zeros((self.n, self.n))
        Gam = np.zeros((self.n,) * 4)
        for wk, j in zip(wts, pick):
            g1, g2 = self.rdms(V[:, j])
            gam += wk * g1
            Gam += wk * g2
        return w[pick], V[:, pick], gam, Gam


# ------------------------------------------------------------- J/K providers
def gpu_jk(ctx, prec="fp64"):
    """The GPU build as a J/K provider: one batched call per list of items."""
    def jk(items):
        if len(items) == 1:
            D, C, nocc = items[0]
            return [ctx.build(D, C, nocc, prec=prec)]
        return ctx.build_batch(items, prec=prec)
    return jk


# ------------------------------------------------------------------ CASSCF
class CASSCF:
    """SA-CASSCF(n_elec_act, n_act) over a closed-shell core of n_core
    orbitals. B_packed is the fitted B (naux x nbf(nbf+1)/2, as
    JKContext.get_B returns it), used for the active integrals only; every
    J/K goes through `jk`."""

    def __init__(self, mol, jk, B_packed, n_core, n_act, n_alpha, n_beta, nstates=1, mult=1, weights=None,
                 H=None):
        self.mol, self.jk = mol, jk
        self.nc, self.na = n_core, n_act
        self.space = ActiveSpace(n_act, n_alpha, n_beta)
        self.nstates, self.mult, self.weights = nstates, mult, weights
        n = mol.nbf
        self.B = unpack_sym(np.asarray(B_packed), n)  # naux x n x n
        if H is None:
            _, T, V = mol.one_electron()
            H = T + V
        self.H = H
        self._guess = None

    def natural_factor(self, gam, Ca):
        """Rows NO_k sqrt(n_k / 2) of gamma's natural orbitals (nat_orb_calc
        with renorm = 1, 'T': negative occupations set to zero)."""
        occ, U = np.linalg.eigh(gam)
        occ = np.maximum(occ, 0.0)
        return (np.sqrt(0.5 * occ)[:, None] * (Ca @ U).T)

    def pair_fock(self, C, gam):
        """F_I and F_I + F_A from one batched call of NOPT's pair."""
        nc, na = self.nc, self.na
        Cc, Ca = C[:, :nc], C[:, nc:nc + na]
        Dc = Cc @ Cc.T
        Da = Ca @ gam @ Ca.T
        X = np.vstack([Cc.T, self.natural_factor(gam, Ca)])
        (J1, K1), (J2, K2) = self.jk([(Dc, np.ascontiguousarray(Cc.T), nc),
                                      (2.0 * Dc + Da, np.ascontiguousarray(X), nc + na)])
        return self.H + 2.0 * J1 - K1, self.H + J2 - K2

    def evaluate(self, C):
        """Energy, gradient (non-redundant block list), the state energies,
        and the pieces of the step at orbitals C."""
        nc, na, mol = self.nc, self.na, self.mol
        nmo = C.shape[1]
        Cc, Ca = C[:, :nc], C[:, nc:nc + na]
        Dc = Cc @ Cc.T
        (J1, K1), = self.jk([(Dc, np.ascontiguousarray(Cc.T), nc)])
        FI = self.H + 2.0 * J1 - K1
        e_core = float(np.sum(Dc * (self.H + FI))) + mol.e_nuc
        Bpu = np.einsum("Qmn,mp,nu->Qpu", self.B, C, Ca, optimize=True)  # (Q|pu)
        Baa = Bpu[:, nc:nc + na, :]
        g_aaaa = np.einsum("Qtu,Qvw->tuvw", Baa, Baa, optimize=True)
        FI_mo = C.T @ FI @ C
        h_act = FI_mo[nc:nc + na, nc:nc + na]
        eci, vecs, gam, Gam = self.space.solve(h_act, g_aaaa, self.nstates, self.mult, self.weights,
                                               guess=self._guess)
        self._guess = vecs  # Lanczos starts from the last ground state
        E = e_core + float(np.sum(h_act * gam)) + 0.5 * float(np.sum(g_aaaa * Gam))
        # total Fock from the second build of the pair
        Da = Ca @ gam @ Ca.T
        X = np.vstack([Cc.T, self.natural_factor(gam, Ca)])
        (J2, K2), = self.jk([(2.0 * Dc + Da, np.ascontiguousarray(X), nc + na)])
        Ft = self.H + J2 - K2
        Ft_mo = C.T @ Ft @ C
        # generalised Fock
        G = np.zeros((nmo, nmo))
        G[:nc, :] = 2.0 * Ft_mo[:, :nc].T
        g_paaa = np.einsum("Qpu,Qvw->puvw", Bpu, Baa, optimize=True)  # (pu|vw)
        G[nc:nc + na, :] = gam @ FI_mo[:, nc:nc + na].T + np.einsum("tuvw,quvw->tq", Gam, g_paaa)
        grad = 2.0 * (G.T - G)
        return dict(E=E, grad=grad, e_states=eci + e_core, gamma=gam, Gamma=Gam, FI=FI, Ft=Ft,
                    Ft_mo=Ft_mo, vecs=vecs)

    def mask(self, nmo):
        """Non-redundant rotations kappa_pq, p > q: active-core, virtual-core,
        virtual-active."""
        nc, na = self.nc, self.na
        m = np.zeros((nmo, nmo), dtype=bool)
        m[nc:nc + na, :nc] = True
        m[nc + na:, :nc + na] = True
        return m

    def run(self, C0, g_conv=1e-7, e_conv=1e-11, max_it=300, memory=10, trust=0.3, verbose=False):
        """L-BFGS over the non-redundant rotations from orbitals C0 (columns,
        orthonormal in the overlap metric)."""
        C = np.array(C0, dtype=np.float64)
        nmo = C.shape[1]
        msk = self.mask(nmo)
        r = self.evaluate(C)
        S_hist, Y_hist, hist = [], [], []
        E_prev = None
        for it in range(max_it):
            g = r["grad"][msk]
            gmax = float(np.abs(g).max())
            hist.append((it, r["E"], gmax))
            if verbose:
                print(f"CAS {it} E={r['E']:.12f} |g|max={gmax:.3e} trust={trust:.3f}")
            if gmax < g_conv and E_prev is not None and abs(r["E"] - E_prev) < e_conv:
                r.update(C=C, iters=it, gmax=gmax, history=hist, converged=True)
                return r
            f = np.diag(r["Ft_mo"])
            P, Qi = np.nonzero(msk)
            hdiag = np.maximum(np.abs(4.0 * (f[P] - f[Qi])), 0.2)
            q = g.copy()
            saved = []
            for s_, y_ in zip(reversed(S_hist), reversed(Y_hist)):
                rho = 1.0 / float(s_ @ y_)
                a = rho * float(s_ @ q)
                saved.append((a, rho, s_, y_))
                q -= a * y_
            step = q / hdiag
            for a, rho, s_, y_ in reversed(saved):
                step += (a - rho * float(y_ @ step)) * s_
            step = -step
            if float(step @ g) >= 0.0:
                step = -g / hdiag
                S_hist.clear()
                Y_hist.clear()
            big = float(np.abs(step).max())
            if big > trust:
                step *= trust / big
            kap = np.zeros((nmo, nmo))
            kap[msk] = step
            kap = kap - kap.T
            Cn = C @ sl.expm(kap)
            rn = self.evaluate(Cn)
            if rn["E"] > r["E"] + 1e-12 * max(1.0, abs(r["E"])):
                trust = max(0.5 * trust, 1e-4)
                S_hist.clear()
                Y_hist.clear()
                continue
            y = rn["grad"][msk] - g
            if float(step @ y) > 1e-14:
                S_hist.append(step)
                Y_hist.append(y)
                if len(S_hist) > memory:
                    S_hist.pop(0)
                    Y_hist.pop(0)
            if big <= trust:
                trust = min(2.0 * trust, 1.0)
            E_prev = r["E"]
            C, r = Cn, rn
        r.update(C=C, iters=max_it, gmax=float(np.abs(r["grad"][msk]).max()), history=hist, converged=False)
        return r
