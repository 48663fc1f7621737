"""ctypes binding of libnoptjk.so (include/nopt_jk.h, include/nopt_ints.h).

The library is built in-tree by ``python -m nopt_amd.build``; loading fails
loudly when it is missing, there is no fallback path."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnoptjk.so")
BASIS_DIR = os.path.join(HERE, "basis")

NOPT_OK, NOPT_EINVAL, NOPT_ENOMEM, NOPT_ECHOL, NOPT_EHIP, NOPT_EIO, NOPT_ESTATE = range(7)
MAX_BATCH = 4  # NOPT_MAX_BATCH
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16x6": 3, "int8": 4}

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_long_p = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class NoptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes); the symbols every build of the library exports
HOST_SYMBOLS = [
    ("nopt_mol_create", ctypes.c_int, [ctypes.c_int, _c_int_p, _c_double_p, ctypes.c_char_p,
                                       ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("nopt_mol_destroy", None, [_vp]),
    ("nopt_mol_dims", ctypes.c_int, [_vp, _c_long_p, _c_long_p, _c_long_p, _c_double_p]),
    ("nopt_mol_function_l", ctypes.c_int, [_vp, _c_int_p]),
    ("nopt_mol_one_electron", ctypes.c_int, [_vp, _c_double_p, _c_double_p, _c_double_p]),
    ("nopt_mol_metric", ctypes.c_int, [_vp, _c_double_p]),
    ("nopt_mol_eri3", ctypes.c_int, [_vp, _c_double_p, ctypes.c_int]),
    ("nopt_set_host_threads", None, [ctypes.c_int]),
    ("nopt_last_error", ctypes.c_char_p, []),
]
GPU_SYMBOLS = [
    ("nopt_jk_create", ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.POINTER(_vp)]),
    ("nopt_jk_create_ex", ctypes.c_int, [ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                         ctypes.POINTER(_vp)]),
    ("nopt_jk_get_tuning", ctypes.c_int, [_vp, _c_int_p]),
    ("nopt_jk_set_tuning", ctypes.c_int, [_vp, _c_int_p]),
    ("nopt_jk_autotune", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long, ctypes.c_int, ctypes.c_int, _c_double_p, _vp]),
    ("nopt_jk_destroy", None, [_vp]),
    ("nopt_jk_set_B", ctypes.c_int, [_vp, _c_double_p, _c_double_p]),
    ("nopt_jk_set_B_fitted", ctypes.c_int, [_vp, _c_double_p]),
    ("nopt_jk_get_B", ctypes.c_int, [_vp, _c_double_p]),
    ("nopt_jk_device_bytes", ctypes.c_long, [_vp]),
    ("nopt_jk_zero_blocks", ctypes.c_int, [_vp, _c_double_p]),
    ("nopt_jk_active_ints", ctypes.c_int, [_vp, _c_double_p, _c_double_p, ctypes.c_long, _c_double_p]),
    ("nopt_jk_build", ctypes.c_int, [_vp, _c_double_p, _c_double_p, ctypes.c_long, _c_double_p,
                                     _c_double_p, ctypes.c_int]),
    ("nopt_jk_build_device", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long, _vp, _vp, ctypes.c_int, _vp]),
    ("nopt_jk_build_device_static", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long, _vp, _vp, ctypes.c_int, _vp]),
    ("nopt_jk_static_scratch_bytes", ctypes.c_long, []),
    ("nopt_jk_build_batch", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_c_double_p),
                                           ctypes.POINTER(_c_double_p), _c_long_p, ctypes.POINTER(_c_double_p),
                                           ctypes.POINTER(_c_double_p), ctypes.c_int]),
    ("nopt_jk_build_batch_device", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                  _c_long_p, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                  ctypes.c_int, _vp]),
    ("nopt_jk_build_ragged_device", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), _c_int_p, ctypes.POINTER(_vp),
                                                   ctypes.POINTER(_vp), _c_long_p, ctypes.POINTER(_vp),
                                                   ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    ("nopt_ragged_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), _c_int_p, ctypes.POINTER(_vp),
                                          ctypes.POINTER(_vp), _c_long_p, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                          ctypes.POINTER(_vp)]),
    ("nopt_ragged_launch", ctypes.c_int, [_vp, _vp]),
    ("nopt_ragged_launches", ctypes.c_int, [_vp]),
    ("nopt_ragged_destroy", None, [_vp]),
    ("nopt_jk_fock_energy", ctypes.c_int, [_vp, _c_double_p, _c_double_p, _c_double_p, _c_double_p,
                                           ctypes.c_double, ctypes.c_double, _c_double_p, _c_double_p]),
    ("nopt_jk_energy_gradient", ctypes.c_int, [_vp, _c_double_p, _c_double_p, ctypes.c_long, ctypes.c_int,
                                               _c_double_p, _c_double_p, _c_double_p]),
    ("nopt_jk_energy_gradient_device", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long, ctypes.c_int, _vp, _vp, _vp,
                                                      _vp]),
    ("nopt_jk_set_timing", ctypes.c_int, [_vp, ctypes.c_int]),
    ("nopt_jk_read_timing", ctypes.c_int, [_vp, _c_double_p, _c_long_p]),
    ("nopt_jk_stream", _vp, [_vp]),
    ("nopt_dev_malloc", _vp, [ctypes.c_int, ctypes.c_size_t]),
    ("nopt_dev_free", None, [_vp]),
    ("nopt_dev_memcpy", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, ctypes.c_int]),
    ("nopt_dev_synchronize", ctypes.c_int, [ctypes.c_int]),
    ("nopt_graph_capture_begin", ctypes.c_int, [ctypes.c_int, _vp]),
    ("nopt_graph_capture_end", ctypes.c_int, [ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    ("nopt_graph_launch", ctypes.c_int, [_vp, _vp]),
    ("nopt_graph_destroy", None, [_vp]),
    # NOPT-named RI entry points (include/nopt_ri.h)
    ("RI_gpu_attach", ctypes.c_int, [_c_double_p, _c_double_p, ctypes.c_long, ctypes.c_long, ctypes.c_int]),
    ("RI_gpu_attach_molecule", ctypes.c_int, [ctypes.c_int, _c_int_p, _c_double_p, ctypes.c_char_p,
                                              ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    ("RI_gpu_set_precision", ctypes.c_int, [ctypes.c_int]),
    ("RI_core_realloc", ctypes.c_int, [ctypes.c_long, ctypes.c_long]),
    ("regen_RI_AA_release", ctypes.c_int, []),
    ("DM_to_F_transform_RI", ctypes.c_int, [_c_double_p, _c_double_p, _c_double_p, _c_double_p,
                                            ctypes.c_long, ctypes.c_long]),
    ("calc_2el_MO_INTS_RI", ctypes.c_int, [ctypes.c_long, _c_double_p, _c_double_p, _c_double_p, _c_double_p,
                                           _c_double_p, _c_double_p, _c_double_p, ctypes.c_long, ctypes.c_long,
                                           ctypes.c_long]),
]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m nopt_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in HOST_SYMBOLS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        for name, res, args in GPU_SYMBOLS:
            f = getattr(L, name, None)
            if f is not None:
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != NOPT_OK:
        raise NoptError(code, lib().nopt_last_error().decode())


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_c_double_p)


def basis_path(name):
    p = os.path.join(BASIS_DIR, name if name.endswith(".gbs") else name + ".gbs")
    return p
