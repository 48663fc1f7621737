"""The C-ABI library loads and exports every entry point that include/*.h
declares (no GPU calls). CPU only."""
import ctypes
import glob
import os
import re

import pytest

from nopt_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # every function declaration: a line that opens with its return type
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b(\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    assert "nopt_jk_build" in names and "nopt_mol_eri3" in names


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports(name):
    lib = ctypes.CDLL(L.LIB_PATH)
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_error_path_without_gpu_work():
    """A bad argument is reported through nopt_last_error, not a crash."""
    lib = L.lib()
    rc = lib.nopt_mol_create(0, None, None, None, None, 1, None)
    assert rc == L.NOPT_EINVAL
    assert b"null" in lib.nopt_last_error()
