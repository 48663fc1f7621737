"""The C-ABI library loads and exports every entry point that include/*.h
declares (no GPU calls). CPU only."""
import ctypes
import glob
import os
import re
import subprocess

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


# C++ (mangled) names of NOPT's include/RI.h:10-16, exported beside the C ones
RI_H_CXX = ["_Z15RI_core_reallocll", "_Z20DM_to_F_transform_RIPdS_S_S_ll",
            "_Z19calc_2el_MO_INTS_RIlPdS_S_S_S_S_S_lll"]


@pytest.mark.parametrize("name", RI_H_CXX)
def test_library_exports_ri_h_cxx_names(name):
    lib = ctypes.CDLL(L.LIB_PATH)
    assert hasattr(lib, name)


def test_cxx_caller_of_ri_h_links_unchanged(tmp_path):
    """A C++ translation unit that declares the three functions exactly as
    NOPT's RI.h does compiles, links against libnoptjk.so and runs (no RI
    context: each call returns an error before any device work)."""
    exe = str(tmp_path / "ri_h_caller")
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", os.path.join(ROOT, "tests", "cxx", "ri_h_caller.cpp"), "-o", exe,
                           "-L" + libdir, "-l:libnoptjk.so", "-Wl,-rpath," + libdir])
    assert subprocess.run([exe], timeout=60).returncode == 0
