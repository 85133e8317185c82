"""The C ABI library loads on a machine without a GPU and exports every entry
point include/mcx.h declares; the Python binding covers all of them.  No
compute call is made here."""
import ctypes
import os
import re

import pytest

from mc_cpp_amd import mcx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mcx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for s in ("mcx_create", "mcx_energy", "mcx_update_list", "mcx_get_list", "mcx_delta_energy", "mcx_mc_run",
              "mcx_nonbonded_grad"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(mcx.LIBPATH):
        pytest.skip("libmcx.so not built")
    lib = ctypes.CDLL(mcx.LIBPATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def fortran_symbols():
    text = open(os.path.join(ROOT, "include", "mcx_fortran.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)) - {"defined"})


def test_library_exports_the_fortran_interface():
    """include/mcx_fortran.h: the names of src/FORTRAN_interface.cpp:23-105."""
    syms = fortran_symbols()
    assert {"set_pbc_", "set_ens_", "set_ff_and_cor_", "get_energy_", "set_coords_", "get_mass_", "clean_"} <= set(syms)
    if not os.path.exists(mcx.LIBPATH):
        pytest.skip("libmcx.so not built")
    lib = ctypes.CDLL(mcx.LIBPATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_fortran_interface_reports_errors_without_exit(tmp_path):
    """A missing coordinates file is reported, not exit()ed (the reference
    exits, src/IO_MDBAS.cpp); no device call is made."""
    if not os.path.exists(mcx.LIBPATH):
        pytest.skip("libmcx.so not built")
    lib = ctypes.CDLL(mcx.LIBPATH)
    lib.mcx_f_last_error.restype = ctypes.c_char_p
    d = ctypes.c_double
    lib.set_pbc_(b"cubic\0", ctypes.byref(d(30.0)), ctypes.byref(d(30.0)), ctypes.byref(d(30.0)),
                 ctypes.byref(d(90.0)), ctypes.byref(d(90.0)), ctypes.byref(d(90.0)))
    lib.set_ens_(ctypes.byref(ctypes.c_int(12)), ctypes.byref(d(300.0)))
    assert lib.mcx_f_last_error() == b""
    lib.set_ff_and_cor_(b"full\0", ctypes.byref(d(12.0)), ctypes.byref(d(10.0)), ctypes.byref(d(2.0)),
                        str(tmp_path / "absent.dat").encode() + b"\0", str(tmp_path / "absent.cor").encode() + b"\0")
    assert b"absent.cor" in lib.mcx_f_last_error()
    lib.clean_()


def test_binding_covers_every_declared_symbol():
    assert sorted(mcx.SIGNATURES) == declared_symbols()


def test_missing_library_fails_loudly(tmp_path):
    saved = mcx._lib
    mcx._lib = None
    try:
        with pytest.raises(RuntimeError):
            mcx.load_library(str(tmp_path / "absent.so"))
    finally:
        mcx._lib = saved


def test_ffield_adapter_compiles_against_the_reference_headers(tmp_path):
    """include/mcx_ffield_adapter.hpp is the FField subclass a maintainer adds
    to the reference tree; it must compile against the reference's headers."""
    import shutil
    import subprocess
    ref_inc = "/root/reference/include"
    if not os.path.isdir(ref_inc) or shutil.which("g++") is None:
        pytest.skip("reference headers not present")
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", "-include", "cmath", "-I", ref_inc,
                        "-I", os.path.join(ROOT, "include"), "-x", "c++",
                        os.path.join(ROOT, "include", "mcx_ffield_adapter.hpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
