"""ORACLE -- test infrastructure only. Imported by tests/, by
``__graft_entry__.smoke()`` and by the CPU-baseline leg of bench.py, always as
the checker and never as the thing measured or shipped; the product
(``nopt_amd``) does not import it.

FP64 CPU restatement of NOPT's closed-shell RI-RHF around the RI-J/K build:
  * basis sets: Psi4 .gbs parsing (src/basis_lib_read.cpp:57-247); the
    orbital basis is Cartesian, as NOPT reads it (pure=false,
    src/inp_par_read.cpp:2550-2552), the auxiliary basis spherical;
  * integrals: Obara-Saika in oracle_ints.c (libint2's published algorithm);
  * D5: the MOs are confined to the spherical (harmonic) subspace of every
    Cartesian shell, which is what NOPT's T_D5 does with D5=1
    (src/molecule.cpp:4551-4839, 4010-4056). The subspace is built here as
    the null space of the Laplacian on degree-l monomials, independently of
    any solid-harmonic formula; only its span matters for the energy;
  * J/K: oracle_jk.c, NOPT's DM_to_F_transform_RI (src/RI.cpp:115-176) with
    V^-1/2 from S05_calc (src/matr.cpp:372-401);
  * Fock and energy: F = H + 2J - K, E = sum DF + sum DH + V_nuc
    (src/molecule.cpp:4407-4438, src/SCF.cpp:134, src/doCI_matr.cpp:38-47);
    V_nuc as E_nuc_calc (src/molecule.cpp:2236-2249); Angstrom -> bohr
    factor 1.889725989 (src/molecule.cpp:472).
NOPT converges with SOSCF/L-BFGS (src/SCF.cpp:108-170); the oracle converges
the same fixed point with DIIS, since only the converged energy is compared.

Parity pins (tests/test_oracle.py): the RI-RHF energies of C2H4, C6H8 and
guanine printed in NOPT's example logs; closed forms of s-type integrals and
the Boys function; the factored J/K against the explicit four-index RI
contraction on small molecules.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
ANG2BOHR = 1.889725989
SYMBOL_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
AM = {"S": 0, "P": 1, "D": 2, "F": 3, "G": 4, "H": 5, "I": 6}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-B", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_boys.restype = ctypes.c_double
        L.oracle_boys.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_s05.restype = ctypes.c_int
        L.oracle_s05.argtypes = [_dp, _dp, _dp, ctypes.c_int]
        L.oracle_jk_nopt.restype = ctypes.c_int
        L.oracle_jk_nopt.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                     _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_prepare_b.restype = None
        L.oracle_prepare_b.argtypes = [_dp, _dp, _dp, ctypes.c_long, ctypes.c_long]
        L.oracle_jk_nopt_prepared.restype = ctypes.c_int
        L.oracle_jk_nopt_prepared.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                              _dp, _dp, _dp]
        _lib = L
        L.oracle_set_threads(int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1)))
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


# ------------------------------------------------------------------ basis --
def read_gbs(path):
    """{symbol: [(l, [alpha], [coef]), ...]} from a Psi4 .gbs file."""
    out, cur, sym = {}, None, None
    lines = [ln.split("!")[0].strip() for ln in open(path)]
    lines = [ln for ln in lines if ln]
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln in ("****", "spherical", "cartesian"):
            i += 1
            continue
        f = ln.split()
        if len(f) == 2 and f[1] == "0" and sym is None:
            sym, cur = f[0], []
            i += 1
            while i < len(lines) and lines[i] != "****":
                g = lines[i].split()
                am, npr, scale = g[0].upper(), int(g[1]), float(g[2])
                rows = [lines[i + 1 + k].replace("D", "E").split() for k in range(npr)]
                alpha = [float(r[0]) * scale * scale for r in rows]
                if am == "SP":
                    cur.append((0, alpha, [float(r[1]) for r in rows]))
                    cur.append((1, alpha, [float(r[2]) for r in rows]))
                else:
                    cur.append((AM[am], alpha, [float(r[1]) for r in rows]))
                i += npr + 1
            out[sym.capitalize()] = cur
            sym = None
        i += 1
    return out


class Shells:
    """Flat shell arrays for oracle_ints.c."""

    def __init__(self, lib_basis, symbols, xyz):
        ls, cen, npr, off, al, co = [], [], [], [], [], []
        self.atom = []
        for a, (s, r) in enumerate(zip(symbols, xyz)):
            for (l, alpha, coef) in lib_basis[s.capitalize()]:
                ls.append(l)
                cen.extend(r)
                npr.append(len(alpha))
                off.append(len(al))
                al.extend(alpha)
                co.extend(coef)
                self.atom.append(a)
        self.l = np.array(ls, dtype=np.int32)
        self.center = np.array(cen, dtype=np.float64)
        self.np = np.array(npr, dtype=np.int32)
        self.off = np.array(off, dtype=np.int32)
        self.alpha = np.array(al, dtype=np.float64)
        self.coef = np.array(co, dtype=np.float64)
        self.n = len(ls)

    def args(self):
        return (self.n, _i(self.l), _d(self.center), _i(self.np), _i(self.off), _d(self.alpha), _d(self.coef))

    def ncart(self):
        return int(sum((l + 1) * (l + 2) // 2 for l in self.l))

    def nsph(self):
        return int(sum(2 * l + 1 for l in self.l))


def monomials(l):
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


def _dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return float(r)


def harmonic_basis(l):
    """(2l+1) x ncart(l): rows span ker(Laplacian) on degree-l monomials and
    are orthonormal in the Gaussian metric of x^l-normalised components."""
    mon = monomials(l)
    nc = len(mon)
    if l < 2:
        return np.eye(nc)
    low = {m: k for k, m in enumerate(monomials(l - 2))}
    lap = np.zeros((len(low), nc))
    for k, (a, b, c) in enumerate(mon):
        for d, e in enumerate((a, b, c)):
            if e >= 2:
                t = [a, b, c]
                t[d] -= 2
                lap[low[tuple(t)], k] += e * (e - 1)
    _, s, vt = np.linalg.svd(lap)
    null = vt[len(low):]  # rank of the Laplacian map is ncart(l-2)
    G = np.zeros((nc, nc))
    for i, (a, b, c) in enumerate(mon):
        for j, (d, e, f) in enumerate(mon):
            if (a + d) % 2 or (b + e) % 2 or (c + f) % 2:
                continue
            G[i, j] = _dfact(a + d - 1) * _dfact(b + e - 1) * _dfact(c + f - 1) / _dfact(2 * l - 1)
    M = null @ G @ null.T
    w, U = np.linalg.eigh(M)
    return (U @ np.diag(w ** -0.5) @ U.T) @ null


def harmonic_tf(lmax):
    return np.ascontiguousarray(np.concatenate([harmonic_basis(l).ravel() for l in range(lmax + 1)]))


# ------------------------------------------------------------- molecule ----
class OracleMolecule:
    def __init__(self, symbols, xyz_bohr, basis_path, aux_path):
        self.symbols = list(symbols)
        self.xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.Z = np.array([SYMBOL_Z[s.capitalize()] for s in self.symbols], dtype=np.float64)
        self.orb = Shells(read_gbs(basis_path), self.symbols, self.xyz)
        self.aux = Shells(read_gbs(aux_path), self.symbols, self.xyz)
        self.n_ao = self.orb.ncart()
        self.naux = self.aux.nsph()
        self.nelec = int(self.Z.sum())
        self.nocc = self.nelec // 2
        self.tf = harmonic_tf(int(max(self.aux.l.max(), self.orb.l.max())))
        n = len(self.Z)
        self.e_nuc = sum(self.Z[i] * self.Z[j] / np.linalg.norm(self.xyz[i] - self.xyz[j])
                         for i in range(n) for j in range(i + 1, n))

    @classmethod
    def from_angstrom(cls, symbols, xyz_ang, basis_path, aux_path):
        return cls(symbols, np.asarray(xyz_ang, dtype=np.float64) * ANG2BOHR, basis_path, aux_path)

    def one_electron(self):
        n = self.n_ao
        S, T, V = (np.zeros((n, n)) for _ in range(3))
        xyz = np.ascontiguousarray(self.xyz.ravel())
        lib().oracle_one_electron(*self.orb.args(), len(self.Z), _d(self.Z), _d(xyz), _d(S), _d(T), _d(V))
        return S, T, V

    def metric(self):
        V = np.zeros((self.naux, self.naux))
        lib().oracle_metric(*self.aux.args(), _d(self.tf), _d(V))
        return V

    def eri3_nopt(self):
        out = np.zeros((self.n_ao, self.n_ao, self.naux))
        lib().oracle_eri3(*self.orb.args(), *self.aux.args(), _d(self.tf), _d(out))
        return out

    def d5(self):
        """n_sph x n_ao block-diagonal map onto the harmonic subspace."""
        rows = []
        col = 0
        T = np.zeros((self.orb.nsph(), self.n_ao))
        r = 0
        for l in self.orb.l:
            h = harmonic_basis(int(l))
            T[r:r + h.shape[0], col:col + h.shape[1]] = h
            r += h.shape[0]
            col += h.shape[1]
        return T


# ------------------------------------------------------------------- J/K ---
def s05(V):
    n = V.shape[0]
    A = np.ascontiguousarray(V.copy())
    M = np.zeros((n, n))
    P = np.zeros((n, n))
    dropped = lib().oracle_s05(_d(A), _d(M), _d(P), n)
    return M, P, dropped


class NoptJK:
    """DM_to_F_transform_RI with NOPT's buffers (src/RI.cpp:16-32, 115-176)."""

    def __init__(self, RI_AA, Vm05, n_c):
        self.RI_AA = np.ascontiguousarray(RI_AA)
        self.n_ao = RI_AA.shape[0]
        self.naux = RI_AA.shape[2]
        self.Vm05 = np.ascontiguousarray(Vm05)
        self.realloc(n_c)
        self.RI_J = np.zeros(self.naux)
        self.RI_J2 = np.zeros(self.naux)

    def realloc(self, n_c):
        """RI_core_realloc (src/RI.cpp:104-113)."""
        self.n_c = n_c
        self.RI_CA = np.zeros(n_c * self.n_ao * self.naux)
        self.RI_CA2 = np.zeros(n_c * self.n_ao * self.naux)

    def __call__(self, DM, MO, n_c):
        if n_c > self.n_c:
            raise ValueError("n_c exceeds the scratch size; call realloc first")
        n = self.n_ao
        J = np.zeros((n, n))
        K = np.zeros((n, n))
        DM = np.ascontiguousarray(DM, dtype=np.float64)
        MO = np.ascontiguousarray(MO, dtype=np.float64)
        lib().oracle_jk_nopt(_d(J), _d(K), _d(DM), _d(MO), n_c, n, self.naux, _d(self.RI_AA), _d(self.Vm05),
                             _d(self.RI_CA), _d(self.RI_CA2), _d(self.RI_J), _d(self.RI_J2))
        return J, K


class NoptJKPrepared:
    """The same build over RI_B = RI_AA.V^-1/2, formed once per geometry (no
    V^-1/2 products per build: src/RI.cpp:120-171 without :128-138 and
    :152-156)."""

    def __init__(self, RI_AA, Vm05, n_c):
        RI_AA = np.ascontiguousarray(RI_AA)
        self.n_ao, self.naux = RI_AA.shape[0], RI_AA.shape[2]
        self.RI_B = np.empty_like(RI_AA)
        lib().oracle_prepare_b(_d(self.RI_B), _d(RI_AA), _d(np.ascontiguousarray(Vm05)), self.n_ao, self.naux)
        self.n_c = n_c
        self.RI_CA = np.zeros(2 * n_c * self.n_ao * self.naux)
        self.RI_J = np.zeros(self.naux)

    def __call__(self, DM, MO, n_c):
        if n_c > self.n_c:
            raise ValueError("n_c exceeds the scratch size")
        n = self.n_ao
        J = np.zeros((n, n))
        K = np.zeros((n, n))
        lib().oracle_jk_nopt_prepared(_d(J), _d(K), _d(np.ascontiguousarray(DM, dtype=np.float64)),
                                      _d(np.ascontiguousarray(MO, dtype=np.float64)), n_c, n, self.naux,
                                      _d(self.RI_B), _d(self.RI_CA), _d(self.RI_J))
        return J, K


# -------------------------------------------------------------------- SCF --
def rhf(mol, e_conv=1e-11, d_conv=1e-9, max_it=100, verbose=False):
    """Closed-shell RI-RHF in NOPT's representation (Cartesian AOs, MOs in the
    D5 subspace). Returns dict(E, C_cart (n_ao x nocc), D, F, H, S, jk)."""
    S, T, V = mol.one_electron()
    H = T + V
    Vm05, _, dropped = s05(mol.metric())
    jk = NoptJK(mol.eri3_nopt(), Vm05, mol.nocc)
    Td = mol.d5()
    Ss = Td @ S @ Td.T
    s, U = np.linalg.eigh(Ss)
    X = U @ np.diag(s ** -0.5) @ U.T
    e, Cp = np.linalg.eigh(X.T @ (Td @ H @ Td.T) @ X)
    Cs = X @ Cp
    E_old, dF, dE = 0.0, [], []
    for it in range(max_it):
        Cc = Td.T @ Cs[:, :mol.nocc]  # n_ao x nocc
        D = Cc @ Cc.T
        J, K = jk(D, Cc.T, mol.nocc)
        F = H + 2 * J - K
        E = float(np.sum(D * F) + np.sum(D * H) + mol.e_nuc)
        Fs = Td @ F @ Td.T
        Ds = Cs[:, :mol.nocc] @ Cs[:, :mol.nocc].T
        err = Fs @ Ds @ Ss - Ss @ Ds @ Fs
        dF.append(Fs)
        dE.append(err)
        if len(dF) > 10:
            dF.pop(0)
            dE.pop(0)
        if verbose:
            print(it, "%.12f" % E, np.abs(err).max())
        if abs(E - E_old) < e_conv and np.abs(err).max() < d_conv:
            break
        E_old = E
        nd = len(dF)
        Bm = -np.ones((nd + 1, nd + 1))
        Bm[-1, -1] = 0
        for i in range(nd):
            for j in range(nd):
                Bm[i, j] = np.sum(dE[i] * dE[j])
        rhs = np.zeros(nd + 1)
        rhs[-1] = -1
        cf = np.linalg.lstsq(Bm, rhs, rcond=None)[0][:-1]
        Fd = sum(c * f for c, f in zip(cf, dF))
        e, Cp = np.linalg.eigh(X.T @ Fd @ X)
        Cs = X @ Cp
    return dict(E=E, C=Cc, D=D, F=F, H=H, S=S, J=J, K=K, jk=jk, d5=Td, iters=it, dropped=dropped)


def nopt_mo_ints(RI_AA, Vm05, VEC1, VEC2):
    """The MO_INTS half of calc_2el_MO_INTS_RI (src/RI.cpp:197-226), t = 0:
    RI_aA = VEC1 . RI_AA (:197-201), times V^-1/2 (:203-207), [a][n][Q] ->
    [n][a][Q] (:209-215), RI_aa = VEC2 . RI_aA (:217-221), MO_INTS =
    RI_aa RI_aa^T (:223-226), as an (na, na, na, na) array
    [a2, a1, b2, b1]."""
    n, naux = RI_AA.shape[0], RI_AA.shape[2]
    na = VEC1.shape[0]
    aA = (VEC1 @ RI_AA.reshape(n, n * naux)).reshape(na * n, naux) @ Vm05
    aA = aA.reshape(na, n, naux).transpose(1, 0, 2).reshape(n, na * naux)
    aa = (VEC2 @ aA).reshape(na * na, naux)
    return (aa @ aa.T).reshape(na, na, na, na)


def brute_force_jk(eri3, V, D):
    """Explicit four-index RI contraction, tiny systems only:
    (mn|rs)_RI = sum_PQ (mn|P) V^-1_PQ (Q|rs), formed as W^T W with
    W = L^-1 (P|mn) (V = L L^T); J = (mn|rs) D_rs, K_mn = sum_rs (mr|ns) D_rs."""
    n = eri3.shape[0]
    W = np.linalg.solve(np.linalg.cholesky(V), eri3.reshape(n * n, -1).T)
    g4 = (W.T @ W).reshape(n, n, n, n)
    J = np.einsum("mnrs,rs->mn", g4, D)
    K = np.einsum("mrns,rs->mn", g4, D)
    return J, K
