"""Host-side interface to the RI-J/K build on the GPU (include/nopt_jk.h).

``JKContext`` owns one device context with the fitted B tensor resident in
HBM. Its methods mirror NOPT's RI layer (include/RI.h:6-23):

  set_B(eri3, metric)        gen_RI_AA (src/RI.cpp:34-79), metric folded in
  build(D, C, nocc)          DM_to_F_transform_RI (src/RI.cpp:115-176)
  build_batch(items)         the two builds of a CASSCF step (src/CAS.cpp:696-707)
                             with one pass over B
  fock_energy(H, D, J, K)    molecule::calc_F_AO (src/molecule.cpp:4407-4438)
                             and the RHF energy (src/SCF.cpp:134)
  energy_gradient(H, C)      E_RHF(C) and dE/dC = 4 C F (the orbital gradient
                             g_ia = 4 F_ai of src/converger_2_1.cpp:54-62)

Orbital factors follow NOPT: rows are orbitals (C[i, m]); densities are
normalised to 1 (D = C^T C over the occupied rows, src/SCF.cpp:63-73).
There is no CPU fallback: every call runs the HIP kernels of libnoptjk.so and
fails loudly when the library or the GPU is missing.
"""
import ctypes

import numpy as np

from . import _lib as L


KEEP = {"both": 0, "fp64": 1, "int8": 2}  # NOPT_CTX_DEFAULT, _FP64_ONLY, _INT8_ONLY
COMPACT = 4  # NOPT_CTX_COMPACT


class JKContext:
    """A GPU build context. `keep` selects the forms of B held in HBM: "both"
    (every mode), "fp64" (FP64/FP32/bf16 modes) or "int8" (integer mode).
    `compact` keeps the integer form as the packed triangle alone (the
    half-transform unpacks it; a large system gets it anyway)."""

    def __init__(self, nbf, naux, device=0, keep="both", compact=False):
        lib = L.lib()
        if not hasattr(lib, "nopt_jk_create"):
            raise ImportError("libnoptjk.so was built without the HIP kernels")
        h = ctypes.c_void_p()
        flags = KEEP[keep] | (COMPACT if compact else 0)
        L.check(lib.nopt_jk_create_ex(int(device), int(nbf), int(naux), flags, ctypes.byref(h)))
        self.keep = keep
        self._h = h
        self.nbf, self.naux = int(nbf), int(naux)
        self.npair = self.nbf * (self.nbf + 1) // 2

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().nopt_jk_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def _arr(self, a, shape):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"expected shape {shape}, got {a.shape}")
        return a

    def set_B(self, eri3_packed, metric):
        A = self._arr(eri3_packed, (self.naux, self.npair))
        V = self._arr(metric, (self.naux, self.naux))
        L.check(L.lib().nopt_jk_set_B(self._h, L.dptr(A), L.dptr(V)))

    def set_B_fitted(self, B_packed):
        B = self._arr(B_packed, (self.naux, self.npair))
        L.check(L.lib().nopt_jk_set_B_fitted(self._h, L.dptr(B)))

    def get_B(self):
        out = np.empty((self.naux, self.npair))
        L.check(L.lib().nopt_jk_get_B(self._h, L.dptr(out)))
        return out

    def active_ints(self, C1, C2=None):
        """(t u|v w) over the fitted B for active orbitals as rows of C1 and C2
        (na x nbf; C2 = C1 by default): ints[t2, t1, u2, u1] = sum_Q
        R_Q[t2, t1] R_Q[u2, u1], R_Q[t2, t1] = C1[t1] B_Q C2[t2]^T (the MO_INTS
        of NOPT's calc_2el_MO_INTS_RI)."""
        n = self.nbf
        C1 = np.ascontiguousarray(C1, dtype=np.float64)
        C2 = C1 if C2 is None else np.ascontiguousarray(C2, dtype=np.float64)
        if C1.ndim != 2 or C1.shape[1] != n or C2.shape != C1.shape:
            raise ValueError("C1 and C2 must both be (na, nbf)")
        na = C1.shape[0]
        out = np.empty((na, na, na, na))
        L.check(L.lib().nopt_jk_active_ints(self._h, L.dptr(C1), L.dptr(C2), na, L.dptr(out)))
        return out

    def device_bytes(self):
        return int(L.lib().nopt_jk_device_bytes(self._h))

    def zero_blocks(self):
        """Shares of B's integer forms the integer build reads once the blocks
        of exact zeros are left out: Coulomb segments, Coulomb slots, and
        half-transform slabs (nopt_jk_zero_blocks)."""
        out = np.zeros(3)
        L.check(L.lib().nopt_jk_zero_blocks(self._h, L.dptr(out)))
        return dict(segments=float(out[0]), slots=float(out[1]), slabs=float(out[2]))

    def build(self, D=None, C=None, nocc=None, prec="fp64", J=True, K=True):
        """(J, K) for density D (nbf x nbf) and the first nocc rows of C."""
        n = self.nbf
        p = L.PREC[prec] if isinstance(prec, str) else int(prec)
        Dm = self._arr(D, (n, n)) if D is not None else None
        Cm = None
        if C is not None:
            Cm = np.ascontiguousarray(C, dtype=np.float64)
            if Cm.ndim != 2 or Cm.shape[1] != n:
                raise ValueError("C must be (norb, nbf) with orbitals as rows")
            if nocc is None:
                nocc = Cm.shape[0]
            if nocc > Cm.shape[0]:
                raise ValueError("nocc exceeds the rows of C")
        nocc = int(nocc or 0)
        Jo = np.empty((n, n)) if J else None
        Ko = np.empty((n, n)) if K else None
        lib = L.lib()
        L.check(lib.nopt_jk_build(self._h, L.dptr(Dm) if Dm is not None else None,
                                  L.dptr(Cm) if Cm is not None else None, nocc,
                                  L.dptr(Jo) if Jo is not None else None,
                                  L.dptr(Ko) if Ko is not None else None, p))
        return Jo, Ko

    def build_batch(self, items, prec="fp64", J=True, K=True):
        """Builds for several (D, C, nocc) against the same B in one pass
        (NOPT's inactive + total density pair of a CASSCF step). Each item is
        (D, C, nocc); D or C may be None. Returns [(J, K), ...]."""
        n, nd = self.nbf, len(items)
        if not 1 <= nd <= L.MAX_BATCH:
            raise ValueError(f"1..{L.MAX_BATCH} densities per batch")
        p = L.PREC[prec] if isinstance(prec, str) else int(prec)
        keep, Dp, Cp, no, Jp, Kp, outs = [], [], [], [], [], [], []
        for D, C, nocc in items:
            Dm = self._arr(D, (n, n)) if D is not None else None
            Cm = None
            if C is not None:
                Cm = np.ascontiguousarray(C, dtype=np.float64)
                if Cm.ndim != 2 or Cm.shape[1] != n:
                    raise ValueError("C must be (norb, nbf) with orbitals as rows")
                nocc = Cm.shape[0] if nocc is None else nocc
                if nocc > Cm.shape[0]:
                    raise ValueError("nocc exceeds the rows of C")
            Jo = np.empty((n, n)) if (J and Dm is not None) else None
            Ko = np.empty((n, n)) if K else None
            keep += [Dm, Cm, Jo, Ko]
            Dp.append(L.dptr(Dm) if Dm is not None else None)
            Cp.append(L.dptr(Cm) if Cm is not None else None)
            no.append(int(nocc or 0))
            Jp.append(L.dptr(Jo) if Jo is not None else None)
            Kp.append(L.dptr(Ko) if Ko is not None else None)
            outs.append((Jo, Ko))
        arr = lambda v: (L._c_double_p * nd)(*v)
        L.check(L.lib().nopt_jk_build_batch(self._h, nd, arr(Dp), arr(Cp), (ctypes.c_long * nd)(*no),
                                            arr(Jp), arr(Kp), p))
        return outs

    def build_device(self, dD, dC, nocc, dJ, dK, prec="fp64", stream=0):
        """Device-pointer build (integer addresses), enqueued on `stream`
        (0: the default stream; `self.stream()`: the context's own)."""
        p = L.PREC[prec] if isinstance(prec, str) else int(prec)
        L.check(L.lib().nopt_jk_build_device(self._h, dD, dC, int(nocc), dJ, dK, p, stream or None))

    def build_batch_device(self, items, prec="fp64", stream=0):
        """Device-pointer batch: items of (dD, dC, nocc, dJ, dK) integer
        addresses (0 for none), enqueued on `stream`."""
        nd = len(items)
        p = L.PREC[prec] if isinstance(prec, str) else int(prec)
        col = lambda i: (ctypes.c_void_p * nd)(*[(it[i] or None) for it in items])
        L.check(L.lib().nopt_jk_build_batch_device(self._h, nd, col(0), col(1),
                                                   (ctypes.c_long * nd)(*[int(it[2]) for it in items]),
                                                   col(3), col(4), p, stream or None))

    TUNING_KEYS = ("nsk", "nsj", "nsk8", "nsg", "nsj8", "cplace")

    def tuning(self):
        """Q-split counts and the integer Coulomb placement of the build
        (include/nopt_jk.h, nopt_jk_get_tuning)."""
        a = np.zeros(len(self.TUNING_KEYS), dtype=np.int32)
        L.check(L.lib().nopt_jk_get_tuning(self._h, a.ctypes.data_as(L._c_int_p)))
        return dict(zip(self.TUNING_KEYS, (int(x) for x in a)))

    def set_tuning(self, **kw):
        t = self.tuning()
        t.update(kw)
        a = np.array([t[k] for k in self.TUNING_KEYS], dtype=np.int32)
        L.check(L.lib().nopt_jk_set_tuning(self._h, a.ctypes.data_as(L._c_int_p)))

    def autotune(self, dD, dC, nocc, prec="fp64", reps=3, stream=0):
        """Times the candidate split counts on device buffers (dD, dC: device
        pointers) and keeps the fastest; returns (tuning, ms per build)."""
        ms = ctypes.c_double()
        L.check(L.lib().nopt_jk_autotune(self._h, ctypes.c_void_p(dD), ctypes.c_void_p(dC), int(nocc),
                                         L.PREC[prec], int(reps), ctypes.byref(ms), ctypes.c_void_p(stream)))
        return self.tuning(), ms.value

    def set_timing(self, enable=True):
        L.check(L.lib().nopt_jk_set_timing(self._h, 1 if enable else 0))

    PHASES = ("staging", "half_transform", "gamma", "coulomb_j", "exchange", "j_output")

    def read_timing(self):
        """{phase: summed ms}, builds recorded since the last read."""
        ms = np.zeros(len(self.PHASES))
        nb = ctypes.c_long()
        L.check(L.lib().nopt_jk_read_timing(self._h, L.dptr(ms), ctypes.byref(nb)))
        return dict(zip(self.PHASES, ms.tolist())), nb.value

    def stream(self):
        return L.lib().nopt_jk_stream(self._h)

    def fock_energy(self, H, D, J=None, K=None, cj=2.0, ck=1.0):
        n = self.nbf
        Hm, Dm = self._arr(H, (n, n)), self._arr(D, (n, n))
        Jm = self._arr(J, (n, n)) if J is not None else None
        Km = self._arr(K, (n, n)) if K is not None else None
        F = np.empty((n, n))
        E = ctypes.c_double()
        L.check(L.lib().nopt_jk_fock_energy(self._h, L.dptr(Hm), L.dptr(Dm),
                                            L.dptr(Jm) if Jm is not None else None,
                                            L.dptr(Km) if Km is not None else None,
                                            float(cj), float(ck), L.dptr(F), ctypes.byref(E)))
        return F, E.value

    def energy_gradient(self, H, C, nocc=None, prec="fp64", fock=False):
        """E = sum D(H+F) (no V_nuc) and G = dE/dC = 4 C F (nocc x nbf);
        with fock=True also F."""
        n = self.nbf
        Hm = self._arr(H, (n, n))
        Cm = np.ascontiguousarray(C, dtype=np.float64)
        nocc = int(nocc or Cm.shape[0])
        G = np.empty((nocc, n))
        F = np.empty((n, n)) if fock else None
        E = ctypes.c_double()
        p = L.PREC[prec] if isinstance(prec, str) else int(prec)
        L.check(L.lib().nopt_jk_energy_gradient(self._h, L.dptr(Hm), L.dptr(Cm), nocc, p,
                                                ctypes.byref(E), L.dptr(G), L.dptr(F) if fock else None))
        return (E.value, G, F) if fock else (E.value, G)


class DeviceArray:
    """FP64 buffer in device memory (plumbing for device-pointer builds)."""

    def __init__(self, shape, device=0):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.device = device
        self.ptr = L.lib().nopt_dev_malloc(device, max(self.nbytes, 8))
        if not self.ptr:
            raise MemoryError(L.lib().nopt_last_error().decode())

    @classmethod
    def from_host(cls, a, device=0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.shape, device)
        L.check(L.lib().nopt_dev_memcpy(d.ptr, a.ctypes.data, d.nbytes, 0))
        return d

    def copy_from(self, a):
        """Overwrites the buffer with host array `a` (same size)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        L.check(L.lib().nopt_dev_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0))

    def to_host(self):
        out = np.empty(self.shape)
        L.check(L.lib().nopt_dev_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            L.lib().nopt_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class Graph:
    """Builds captured into one HIP graph (include/nopt_jk.h, nopt_graph_*).

        with Graph.capture(stream) as g:      # stream: a context's stream
            ctx1.build_device(..., stream=stream)
            ctx2.build_device(..., stream=stream)
        g.launch(stream)                       # replays every captured build
    """

    def __init__(self, handle, device):
        self._h, self.device = handle, device

    @classmethod
    def capture(cls, stream, device=0):
        return _Capture(stream, device)

    def launch(self, stream):
        L.check(L.lib().nopt_graph_launch(self._h, stream))

    def close(self):
        if getattr(self, "_h", None):
            L.lib().nopt_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class _Capture:
    def __init__(self, stream, device):
        self.stream, self.device, self.graph = stream, device, None

    def __enter__(self):
        L.check(L.lib().nopt_graph_capture_begin(self.device, self.stream))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        rc = L.lib().nopt_graph_capture_end(self.device, self.stream, ctypes.byref(h))
        if exc_type is None:
            L.check(rc)
            self.graph = Graph(h, self.device)
        elif rc == 0 and h.value:
            L.lib().nopt_graph_destroy(h)
        return False


def build_ragged_device(systems, prec="fp64", stream=0):
    """One call for a ragged batch (include/nopt_jk.h, nopt_jk_build_ragged_device):
    `systems` is a list of (JKContext, items) with items as in
    JKContext.build_batch_device, enqueued on `stream` one system after the
    other once every argument has been checked."""
    nsys = len(systems)
    flat = [it for _, items in systems for it in items]
    n = len(flat)
    col = lambda i: (ctypes.c_void_p * n)(*[(it[i] or None) for it in flat])
    p = L.PREC[prec] if isinstance(prec, str) else int(prec)
    L.check(L.lib().nopt_jk_build_ragged_device(
        nsys, (ctypes.c_void_p * nsys)(*[c.handle.value for c, _ in systems]),
        (ctypes.c_int * nsys)(*[len(items) for _, items in systems]), col(0), col(1),
        (ctypes.c_long * n)(*[int(it[2]) for it in flat]), col(3), col(4), p, stream or None))


class RaggedPlan:
    """A ragged batch of integer builds as work-list launches (include/nopt_jk.h,
    nopt_ragged_create): `systems` as for build_ragged_device. The contexts stay
    pinned until close()."""

    def __init__(self, systems):
        nsys = len(systems)
        flat = [it for _, items in systems for it in items]
        n = len(flat)
        col = lambda i: (ctypes.c_void_p * n)(*[(it[i] or None) for it in flat])
        h = ctypes.c_void_p()
        L.check(L.lib().nopt_ragged_create(
            nsys, (ctypes.c_void_p * nsys)(*[c.handle.value for c, _ in systems]),
            (ctypes.c_int * nsys)(*[len(items) for _, items in systems]), col(0), col(1),
            (ctypes.c_long * n)(*[int(it[2]) for it in flat]), col(3), col(4), ctypes.byref(h)))
        self._h = h

    @property
    def launches(self):
        return int(L.lib().nopt_ragged_launches(self._h))

    def launch(self, stream=0):
        L.check(L.lib().nopt_ragged_launch(self._h, stream or None))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().nopt_ragged_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def synchronize(device=0):
    L.check(L.lib().nopt_dev_synchronize(device))


def context_for(mol, device=0):
    """Context with B formed on the device from the molecule's integrals."""
    ctx = JKContext(mol.nbf, mol.naux, device)
    ctx.set_B(mol.eri3_packed(), mol.metric())
    return ctx
