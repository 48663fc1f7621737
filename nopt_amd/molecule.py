"""Molecules and host integrals over the C-ABI of include/nopt_ints.h."""
import ctypes

import numpy as np

from . import _lib as L

ANG2BOHR = 1.889725989  # NOPT's conversion, src/molecule.cpp:472
SYMBOL_Z = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}


class Molecule:
    """Atoms (bohr) with an orbital and an auxiliary basis."""

    def __init__(self, symbols, xyz_bohr, basis="cc-pvdz", aux="cc-pvdz-jkfit", pure=True):
        self.symbols = list(symbols)
        self.Z = np.array([SYMBOL_Z[s] for s in self.symbols], dtype=np.int32)
        self.xyz = np.ascontiguousarray(np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3))
        self.basis, self.aux, self.pure = basis, aux, pure
        h = ctypes.c_void_p()
        lib = L.lib()
        L.check(lib.nopt_mol_create(len(self.Z), self.Z.ctypes.data_as(L._c_int_p), L.dptr(self.xyz),
                                    L.basis_path(basis).encode(), L.basis_path(aux).encode(),
                                    1 if pure else 0, ctypes.byref(h)))
        self._h = h
        nbf, naux, nel = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
        en = ctypes.c_double()
        L.check(lib.nopt_mol_dims(h, ctypes.byref(nbf), ctypes.byref(naux), ctypes.byref(nel), ctypes.byref(en)))
        self.nbf, self.naux, self.nelec, self.e_nuc = nbf.value, naux.value, nel.value, en.value
        self.nocc = self.nelec // 2

    @classmethod
    def from_angstrom(cls, symbols, xyz_ang, **kw):
        return cls(symbols, np.asarray(xyz_ang, dtype=np.float64) * ANG2BOHR, **kw)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().nopt_mol_destroy(h)
            self._h = None

    def function_l(self):
        out = np.zeros(self.nbf, dtype=np.int32)
        L.check(L.lib().nopt_mol_function_l(self._h, out.ctypes.data_as(L._c_int_p)))
        return out

    def one_electron(self):
        n = self.nbf
        S, T, V = (np.zeros((n, n)) for _ in range(3))
        L.check(L.lib().nopt_mol_one_electron(self._h, L.dptr(S), L.dptr(T), L.dptr(V)))
        return S, T, V

    def metric(self):
        V = np.zeros((self.naux, self.naux))
        L.check(L.lib().nopt_mol_metric(self._h, L.dptr(V)))
        return V

    def eri3_packed(self):
        """(P|mn) as [P][m(m+1)/2+n], m >= n."""
        out = np.empty((self.naux, self.nbf * (self.nbf + 1) // 2))
        L.check(L.lib().nopt_mol_eri3(self._h, L.dptr(out), 0))
        return out

    def eri3_nopt(self):
        """(mn|P) in NOPT's RI_AA layout [(m*nbf+n)*naux+P] (src/RI.cpp:73)."""
        out = np.empty((self.nbf, self.nbf, self.naux))
        L.check(L.lib().nopt_mol_eri3(self._h, L.dptr(out), 1))
        return out


def unpack_sym(packed, n):
    """[..., n(n+1)/2] lower-triangle rows -> [..., n, n] symmetric."""
    lead = packed.shape[:-1]
    out = np.zeros(lead + (n, n), dtype=packed.dtype)
    r, c = np.tril_indices(n)
    out[..., r, c] = packed
    out[..., c, r] = packed
    return out
