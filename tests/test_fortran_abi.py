"""The reference's Fortran/C energy interface (src/FORTRAN_interface.cpp:
21-106) served by libmcx.so (include/mcx_fortran.h), called as a Fortran
program calls it: NUL-terminated strings, arguments by reference, energies in
kcal/mol.  Energies of the reference's own test inputs against the energies
its build computed (tests/golden/golden.json, iu / 418.4), and after
set_coords_ against the FP64 oracle.  Tolerances as for every energy of the
device path: total, elec and vdw 1e-6 relative (of the larger of the
component and the total; the pair kernels' LJ arithmetic is FP32), bonded
terms 1e-9."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from mc_cpp_amd import mcx, systems
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")


def _lib():
    L = C.CDLL(mcx.LIBPATH)
    L.mcx_f_last_error.restype = C.c_char_p
    return L


def _d(v):
    return C.byref(C.c_double(v))


def _setup(L, pbc, box, natom, cut, ff, cor):
    L.set_pbc_(pbc.encode() + b"\0", _d(box[0]), _d(box[1]), _d(box[2]), _d(90.0), _d(90.0), _d(90.0))
    L.set_ens_(C.byref(C.c_int(natom)), _d(300.0))
    L.set_ff_and_cor_(cut.encode() + b"\0", _d(12.0), _d(10.0), _d(2.0), ff.encode() + b"\0", cor.encode() + b"\0")
    assert L.mcx_f_last_error() == b"", L.mcx_f_last_error()


def _check(e, ref):
    tot = abs(ref[0])
    for k in (0, 1, 3, 4):
        assert abs(e[k] - ref[k]) <= 1e-6 * max(abs(ref[k]), tot), (k, e[k], ref[k])
    for k in (2, 5, 6, 7, 8, 9):
        assert abs(e[k] - ref[k]) <= 1e-9 * max(abs(ref[k]), 1e-6), (k, e[k], ref[k])


def _energy(L):
    e = (C.c_double * 10)()
    L.get_energy_(e)
    assert L.mcx_f_last_error() == b"", L.mcx_f_last_error()
    return np.array(e[:])


@pytest.mark.parametrize("name", ["diala_vac", "argon_box"])
def test_fortran_energy_of_reference_inputs(name):
    with open(os.path.join(G, "golden.json")) as f:
        gold = json.load(f)[name]
    ff, cor = {"diala_vac": ("ff.dat", "ala2.cor"), "argon_box": ("forfield.dat", "ar.cor")}[name]
    ff, cor = os.path.join(G, name, ff), os.path.join(G, name, cor)
    cut = "full" if gold["cutmode"] == 0 else "switch"
    L = _lib()
    _setup(L, "cubic", gold["box"], gold["natom"], cut, ff, cor)
    e = _energy(L)
    _check(e, np.array(gold["energy"]) / 418.4)
    # masses as the forcefield file gives them
    s = systems.read_mdbas(cor, ff, gold["box"], cutmode=gold["cutmode"])
    m = np.zeros(gold["natom"])
    L.get_mass_(m.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(m, s.mass)
    # new coordinates: against the oracle's FP64 energy of the same ones,
    # over the list built at set-up (the reference keeps it: set_coords_ is
    # AtomList::setCoords only)
    o = Oracle(s)
    if cut == "switch":
        o.update_list()
    rng = np.random.default_rng(5)
    for arr in (s.x, s.y, s.z):
        arr += rng.uniform(-0.05, 0.05, s.natom)
    P = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.POINTER(C.c_double))
    L.set_coords_(P(s.x), P(s.y), P(s.z))
    e = _energy(L)
    _check(e, o.getE() / 418.4)
    L.clean_()
