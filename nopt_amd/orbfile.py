"""NOPT's binary orbital files (.orb).

Layout (src/binary_files.cpp:248-379, native byte order): int32 13772497
(the format mark, src/binary_files.cpp:16), int32 n_mo, int32 n_ao,
n_mo x n_ao float64 MO coefficients (orbitals as rows, MO[i*n_ao + m]),
n_mo float64 orbital energies, int32 flag; flag 1 is followed by a second
(beta) set of coefficients and energies, flag 0 ends an RHF file.

NOPT writes the coefficients over Cartesian AOs in GAMESS order and
normalisation (molecule::MO_gamess_format, src/molecule.cpp:2545) and turns
them back into libint's order when it reads a file
(molecule::MO_libint_format, src/molecule.cpp:2297). Within a Cartesian shell
of angular momentum l, GAMESS lists the components as below, each normalised
on its own, while the libint order (the order of this package's integrals:
x-power descending, then y-power descending) normalises every component like
x^l; a coefficient over component x^a y^b z^c therefore carries the factor
sqrt((2a-1)!! (2b-1)!! (2c-1)!! / (2l-1)!!) in GAMESS form.
"""
import numpy as np

MAGIC = 13772497

# GAMESS component order of the Cartesian shells, as (a, b, c) powers of x, y, z
GAMESS_ORDER = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2),
        (1, 1, 1)],
    4: [(4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (3, 0, 1), (1, 3, 0), (0, 3, 1), (1, 0, 3), (0, 1, 3),
        (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2)],
}


def _dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def _libint_index(l, a, b):
    """Position of x^a y^b z^(l-a-b) in the libint order of an l shell."""
    i = l - a
    return i * (i + 1) // 2 + (l - a - b)


def _shell_maps(function_l):
    """(position in the AO list, l) of every Cartesian shell: runs of
    functions with the same l, cut into shells of (l+1)(l+2)/2."""
    out, k = [], 0
    fl = np.asarray(function_l)
    while k < len(fl):
        l = int(fl[k])
        nc = (l + 1) * (l + 2) // 2
        if l not in GAMESS_ORDER or k + nc > len(fl) or np.any(fl[k:k + nc] != l):
            raise ValueError(f"function list is not a sequence of Cartesian shells (at function {k}, l={l})")
        out.append((k, l))
        k += nc
    return out


def gamess_to_libint(C, function_l):
    """Coefficients (orbitals as rows) from GAMESS order and normalisation to
    libint's (what NOPT's MO_libint_format does after reading a file)."""
    C = np.asarray(C, dtype=np.float64)
    out = np.empty_like(C)
    for k, l in _shell_maps(function_l):
        for g, (a, b, c) in enumerate(GAMESS_ORDER[l]):
            s = np.sqrt(_dfact(2 * l - 1) / (_dfact(2 * a - 1) * _dfact(2 * b - 1) * _dfact(2 * c - 1)))
            out[..., k + _libint_index(l, a, b)] = C[..., k + g] * s
    return out


def libint_to_gamess(C, function_l):
    """The inverse of gamess_to_libint (NOPT's MO_gamess_format before writing)."""
    C = np.asarray(C, dtype=np.float64)
    out = np.empty_like(C)
    for k, l in _shell_maps(function_l):
        for g, (a, b, c) in enumerate(GAMESS_ORDER[l]):
            s = np.sqrt(_dfact(2 * l - 1) / (_dfact(2 * a - 1) * _dfact(2 * b - 1) * _dfact(2 * c - 1)))
            out[..., k + g] = C[..., k + _libint_index(l, a, b)] / s
    return out


class OrbFile(dict):
    __getattr__ = dict.__getitem__


def read_orb(path):
    """Returns OrbFile(n_mo, n_ao, C (n_mo x n_ao, as stored), eps (n_mo),
    C_beta, eps_beta (None for an RHF file)). Raises ValueError where NOPT's
    reader stops (wrong mark, sizes that do not match the data)."""
    raw = open(path, "rb").read()
    ints = np.frombuffer(raw, dtype=np.int32, count=3) if len(raw) >= 12 else None
    if ints is None or int(ints[0]) != MAGIC:
        raise ValueError(f"{path}: not an orbital file (mark {None if ints is None else int(ints[0])})")
    n_mo, n_ao = int(ints[1]), int(ints[2])
    if n_mo <= 0 or n_ao <= 0:
        raise ValueError(f"{path}: bad sizes n_mo={n_mo} n_ao={n_ao}")
    off = 12
    nc = n_mo * n_ao

    def f64(n):
        nonlocal off
        if off + 8 * n > len(raw):
            raise ValueError(f"{path}: truncated")
        a = np.frombuffer(raw, dtype=np.float64, count=n, offset=off).copy()
        off += 8 * n
        return a

    C = f64(nc).reshape(n_mo, n_ao)
    eps = f64(n_mo)
    if off + 4 > len(raw):
        raise ValueError(f"{path}: truncated")
    flag = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
    off += 4
    Cb = eb = None
    if flag:
        Cb = f64(nc).reshape(n_mo, n_ao)
        eb = f64(n_mo)
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes")
    return OrbFile(n_mo=n_mo, n_ao=n_ao, C=C, eps=eps, C_beta=Cb, eps_beta=eb)


def write_orb(path, C, eps, C_beta=None, eps_beta=None):
    """Writes the layout above (binary_VEC_write, src/binary_files.cpp:347-379)."""
    C = np.ascontiguousarray(C, dtype=np.float64)
    n_mo, n_ao = C.shape
    eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(n_mo)
    with open(path, "wb") as f:
        f.write(np.array([MAGIC, n_mo, n_ao], dtype=np.int32).tobytes())
        f.write(C.tobytes())
        f.write(eps.tobytes())
        if C_beta is not None:
            Cb = np.ascontiguousarray(C_beta, dtype=np.float64).reshape(n_mo, n_ao)
            eb = np.ascontiguousarray(eps_beta, dtype=np.float64).reshape(n_mo)
            f.write(np.array([1], dtype=np.int32).tobytes())
            f.write(Cb.tobytes())
            f.write(eb.tobytes())
        else:
            f.write(np.array([0], dtype=np.int32).tobytes())
