"""ctypes bindings of the checkers: the FP64 restatement (libmc_oracle.so)
and the reference-source harness (_ref/libmcref.so).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmc_oracle.so")
REFLIB = os.path.join(HERE, "_ref", "libmcref.so")
REFMCXLIB = os.path.join(HERE, "_ref", "libmcref_mcx.so")
REF_TREE = "/root/reference"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcSystem(C.Structure):
    _fields_ = [
        ("natom", C.c_int), ("pbtype", C.c_int),
        ("pbx", C.c_double), ("pby", C.c_double), ("pbz", C.c_double),
        ("rpbx", C.c_double), ("rpby", C.c_double), ("rpbz", C.c_double),
        ("cutmode", C.c_int), ("ctoff", C.c_double), ("cton", C.c_double), ("dcut", C.c_double),
        ("temp", C.c_double),
        ("x", _dp), ("y", _dp), ("z", _dp),
        ("charge", _dp), ("eps", _dp), ("sig", _dp), ("eps14", _dp), ("sig14", _dp), ("mass", _dp),
        ("frozen", _ip), ("resid", _ip),
        ("nbond", C.c_int), ("bond_a", _ip), ("bond_b", _ip), ("bond_type", _ip),
        ("bond_k", _dp), ("bond_r0", _dp), ("bond_beta", _dp),
        ("nub", C.c_int), ("ub_a", _ip), ("ub_b", _ip), ("ub_k", _dp), ("ub_r0", _dp),
        ("nangle", C.c_int), ("ang_a", _ip), ("ang_b", _ip), ("ang_c", _ip), ("ang_k", _dp), ("ang_t0", _dp),
        ("ndihe", C.c_int), ("dih_a", _ip), ("dih_b", _ip), ("dih_c", _ip), ("dih_d", _ip), ("dih_type", _ip),
        ("dih_k", _dp), ("dih_phi0", _dp), ("dih_mult", _dp),
        ("nimpr", C.c_int), ("imp_a", _ip), ("imp_b", _ip), ("imp_c", _ip), ("imp_d", _ip), ("imp_type", _ip),
        ("imp_k", _dp), ("imp_phi0", _dp), ("imp_mult", _dp),
        ("exclPair", _ip), ("exclOff", _ip), ("exclIdx", _ip),
        ("nPair14", C.c_int), ("neighList14", _ip), ("pair14_overflow", C.c_int),
        ("neighPair", _ip), ("neighList", _ip), ("cap", C.c_int), ("sizeList", C.c_int),
        ("tot", C.c_double), ("pot", C.c_double), ("kin", C.c_double), ("elec", C.c_double),
        ("vdw", C.c_double), ("bond", C.c_double), ("ang", C.c_double), ("ub", C.c_double),
        ("dihe", C.c_double), ("impr", C.c_double),
    ]


class OrcMoveType(C.Structure):
    _fields_ = [("type", C.c_int), ("mode", C.c_int), ("sele", _ip),
                ("dmax", C.c_double), ("target", C.c_double), ("each", C.c_int)]


class OrcRecord(C.Structure):
    _fields_ = [("mvtyp", _ip), ("mvinst", _ip), ("de", _dp), ("acc", C.POINTER(C.c_ubyte))]


def build(ref: bool = True):
    """Compiles libmc_oracle.so and, when the reference tree is present,
    _ref/libmcref.so (both with -ffp-contract=off)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "libmc_oracle.so"])
    if ref and os.path.isdir(REF_TREE):
        subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "ref", "REF=" + REF_TREE])
        if os.path.exists(os.path.join(HERE, "..", "mc_cpp_amd", "libmcx.so")):
            subprocess.check_call(["make", "-s", "-C", HERE, "ref-mcx", "REF=" + REF_TREE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = C.CDLL(LIB)
        P = C.POINTER(OrcSystem)
        L.orc_build_exclusions.argtypes = [P]
        L.orc_update_verlet_list.argtypes = [P, C.c_int]
        L.orc_update_verlet_list.restype = C.c_longlong
        L.orc_getE.argtypes = [P, _dp]
        L.orc_getE.restype = C.c_double
        L.orc_moving_set_energy.argtypes = [P, C.c_int, C.c_int]
        L.orc_moving_set_energy.restype = C.c_double
        L.orc_moving_set_energy_exact.argtypes = [P, C.c_int, C.c_int]
        L.orc_moving_set_energy_exact.restype = C.c_double
        L.orc_free_lists.argtypes = [P]
        L.orc_set_box.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_mc_run.argtypes = [P, C.POINTER(OrcMoveType), C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int,
                                 C.POINTER(OrcRecord), _dp, C.c_int]
        L.orc_mc_run.restype = C.c_double
        L.orc_mt_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_mt_next.argtypes = [C.c_void_p]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_mt_canonical.argtypes = [C.c_void_p]
        L.orc_mt_canonical.restype = C.c_double
        L.orc_rint.argtypes = [C.c_double]
        L.orc_rint.restype = C.c_double
        L.orc_nonbonded_switch_grad.argtypes = [P] + [_dp] * 6
        L.orc_nonbonded_grad.argtypes = [P, _dp]
        L.orc_set_energy_threads.argtypes = [C.c_int]
        for f in ("orc_nonbonded_full", "orc_nonbonded_switch", "orc_nonbonded14", "orc_nonbonded14_switch",
                  "orc_ebond", "orc_eang", "orc_eub", "orc_edihe", "orc_eimpr"):
            getattr(L, f).argtypes = [P]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """FP64 CPU restatement bound to one System (numpy arrays are shared, so
    coordinate changes made through numpy are seen by the oracle)."""

    def __init__(self, sysm):
        self.L = lib()
        self.sys = sysm
        s = OrcSystem()
        s.natom = sysm.natom
        s.cutmode, s.ctoff, s.cton, s.dcut, s.temp = sysm.cutmode, sysm.ctoff, sysm.cton, sysm.dcut, sysm.temp
        self._keep = []

        def d(a):
            a = np.ascontiguousarray(a, dtype=np.float64); self._keep.append(a); return _p(a, C.c_double)

        def i(a):
            a = np.ascontiguousarray(a, dtype=np.int32); self._keep.append(a); return _p(a, C.c_int)

        for f in ("charge", "eps", "sig", "eps14", "sig14", "mass"):
            setattr(s, f, d(getattr(sysm, f)))
        s.frozen, s.resid = i(sysm.frozen), i(sysm.resid)
        # coordinates are shared with the System object
        assert sysm.x.dtype == np.float64 and sysm.x.flags.c_contiguous
        s.x, s.y, s.z = _p(sysm.x, C.c_double), _p(sysm.y, C.c_double), _p(sysm.z, C.c_double)
        b = sysm.bonds
        s.nbond = len(b["a"]); s.bond_a, s.bond_b, s.bond_type = i(b["a"]), i(b["b"]), i(b["type"])
        s.bond_k, s.bond_r0, s.bond_beta = d(b["k"]), d(b["r0"]), d(b["beta"])
        u = sysm.ub
        s.nub = len(u["a"]); s.ub_a, s.ub_b, s.ub_k, s.ub_r0 = i(u["a"]), i(u["b"]), d(u["k"]), d(u["r0"])
        a = sysm.angles
        s.nangle = len(a["a"]); s.ang_a, s.ang_b, s.ang_c = i(a["a"]), i(a["b"]), i(a["c"])
        s.ang_k, s.ang_t0 = d(a["k"]), d(a["t0"])
        t = sysm.dihedrals
        s.ndihe = len(t["a"])
        s.dih_a, s.dih_b, s.dih_c, s.dih_d, s.dih_type = i(t["a"]), i(t["b"]), i(t["c"]), i(t["d"]), i(t["type"])
        s.dih_k, s.dih_phi0, s.dih_mult = d(t["k"]), d(t["phi0"]), d(t["mult"])
        t = sysm.impropers
        s.nimpr = len(t["a"])
        s.imp_a, s.imp_b, s.imp_c, s.imp_d, s.imp_type = i(t["a"]), i(t["b"]), i(t["c"]), i(t["d"]), i(t["type"])
        s.imp_k, s.imp_phi0, s.imp_mult = d(t["k"]), d(t["phi0"]), d(t["mult"])
        self.s = s
        self.L.orc_set_box(C.byref(s), int(sysm.pbtype), *map(float, sysm.box))
        self.L.orc_build_exclusions(C.byref(s))
        self.nthreads = os.cpu_count() or 1

    def __del__(self):
        try:
            self.L.orc_free_lists(C.byref(self.s))
        except Exception:
            pass

    # ---- lists ----
    def excl_rows(self):
        s, n = self.s, self.sys.natom
        off = np.ctypeslib.as_array(s.exclOff, (n + 1,))
        idx = np.ctypeslib.as_array(s.exclIdx, (max(int(off[-1]), 1),))
        np_ = np.ctypeslib.as_array(s.exclPair, (n,))
        return [idx[off[i]: off[i] + np_[i]].copy() for i in range(n)]

    def excl_csr(self):
        rows = self.excl_rows()
        off = np.zeros(len(rows) + 1, dtype=np.int32)
        off[1:] = np.cumsum([len(r) for r in rows])
        idx = np.concatenate(rows).astype(np.int32) if off[-1] else np.zeros(0, np.int32)
        return off, idx

    def pairs14(self):
        n = self.s.nPair14
        if n == 0:
            return np.zeros((0, 2), dtype=np.int32)
        return np.ctypeslib.as_array(self.s.neighList14, (2 * n,)).reshape(n, 2).copy()

    def update_list(self, nthreads=None):
        return self.L.orc_update_verlet_list(C.byref(self.s), int(nthreads or self.nthreads))

    def list_rows(self):
        s, n = self.s, self.sys.natom
        cnt = np.ctypeslib.as_array(s.neighPair, (n,)).copy()
        flat = np.ctypeslib.as_array(s.neighList, (n * s.cap,)).reshape(n, s.cap)
        return cnt, [flat[i, :cnt[i]].copy() for i in range(n)]

    def list_csr(self):
        cnt, rows = self.list_rows()
        off = np.zeros(len(cnt) + 1, dtype=np.int64)
        off[1:] = np.cumsum(cnt)
        return cnt, off, (np.concatenate(rows).astype(np.int32) if off[-1] else np.zeros(0, np.int32))

    # ---- energies ----
    def getE(self):
        e = np.zeros(10)
        self.L.orc_getE(C.byref(self.s), _p(e, C.c_double))
        return e

    def moving_set_energy(self, first, last):
        return self.L.orc_moving_set_energy(C.byref(self.s), int(first), int(last))

    def moving_set_energy_exact(self, first, last):
        """Every pair the list would hold at the current coordinates."""
        return self.L.orc_moving_set_energy_exact(C.byref(self.s), int(first), int(last))

    def grad_switch(self):
        n = self.sys.natom
        out = [np.zeros(n) for _ in range(6)]
        self.L.orc_nonbonded_switch_grad(C.byref(self.s), *[_p(a, C.c_double) for a in out])
        return out

    def grad_all(self):
        """(8, n): dE/dx, dE/dy, dE/dz, dE/deps, dE/dsig, dE/dq, dE/deps14,
        dE/dsig14 of elec + vdw (list or all pairs, and 1-4 pairs)."""
        n = self.sys.natom
        g = np.zeros(8 * n)
        self.L.orc_nonbonded_grad(C.byref(self.s), _p(g, C.c_double))
        return g.reshape(8, n)

    # ---- Monte Carlo ----
    def mc_run(self, nsteps, seed, mode=0, list_every=50, moves=None, nthreads=None):
        moves = moves if moves is not None else self.sys.moves
        arr = (OrcMoveType * len(moves))()
        keep = []
        for k, m in enumerate(moves):
            arr[k].type, arr[k].mode, arr[k].dmax, arr[k].target, arr[k].each = m.type, m.mode, m.dmax, m.target, m.each
            if m.sele is not None:
                sa = np.ascontiguousarray(m.sele, dtype=np.int32); keep.append(sa)
                arr[k].sele = _p(sa, C.c_int)
        rec_t = np.zeros(nsteps, np.int32); rec_i = np.zeros(nsteps, np.int32)
        rec_de = np.zeros(nsteps); rec_a = np.zeros(nsteps, np.uint8)
        rec = OrcRecord(_p(rec_t, C.c_int), _p(rec_i, C.c_int), _p(rec_de, C.c_double), _p(rec_a, C.c_ubyte))
        dmax = np.zeros(len(moves))
        etot = self.L.orc_mc_run(C.byref(self.s), arr, len(moves), int(nsteps), C.c_uint64(seed), int(mode),
                                 int(list_every), C.byref(rec), _p(dmax, C.c_double), int(nthreads or self.nthreads))
        return dict(etot=etot, mvtyp=rec_t, mvinst=rec_i, de=rec_de, acc=rec_a, dmax=dmax)


class MT19937:
    """The oracle's std::mt19937 + generate_canonical<double,53>."""

    def __init__(self, seed):
        self.buf = C.create_string_buffer(624 * 4 + 8)
        lib().orc_mt_seed(self.buf, C.c_uint64(seed))

    def next(self):
        return lib().orc_mt_next(self.buf)

    def canonical(self):
        return lib().orc_mt_canonical(self.buf)


def rint(x):
    return lib().orc_rint(float(x))


# ---------------------------------------------------------------------------
# Reference-source harness (only where oracle/_ref was built)
# ---------------------------------------------------------------------------

_ref = None


def set_energy_threads(n):
    """Threads of the oracle's switched pair loops (1 = the reference's
    serial order, bit for bit; more = per-thread sums, for CPU timing)."""
    lib().orc_set_energy_threads(int(n))


def ref_available():
    return os.path.exists(REFLIB)


def reflib():
    global _ref
    if _ref is None:
        L = C.CDLL(REFLIB)
        L.refd_output.restype = C.c_char_p
        L.refd_create.restype = C.c_void_p
        L.refd_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_double, C.c_char_p, C.c_double, C.c_double, C.c_double]
        L.refd_destroy.argtypes = [C.c_void_p]
        L.refd_energy.argtypes = [C.c_void_p, _dp]
        L.refd_update_list.argtypes = [C.c_void_p]
        L.refd_get_coords.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.refd_set_coords.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.refd_get_list.argtypes = [C.c_void_p, _ip, _ip, C.c_longlong]
        L.refd_get_list.restype = C.c_longlong
        L.refd_get_excl.argtypes = [C.c_void_p, _ip, _ip, C.c_longlong, _ip, _ip, C.c_int]
        L.refd_get_excl.restype = C.c_longlong
        L.refd_mc.argtypes = [C.c_void_p, C.c_int, _ip, C.POINTER(C.c_char_p), _dp, _dp, _ip, C.c_int,
                              C.c_uint64, _dp]
        _ref = L
    return _ref


class Reference:
    """The reference's own FField_MDBAS / List_nonBonded / MC_metropolis,
    driven through oracle/_ref/libmcref.so on MDBAS files."""

    PBC = {0: b"none", 1: b"cubic", 2: b"orbic"}

    def __init__(self, cor, ff, sysm):
        self.L = reflib()
        self.n = sysm.natom
        cut = b"full" if sysm.cutmode == 0 else b"switch"
        self.h = self.L.refd_create(cor.encode(), ff.encode(), self.PBC[sysm.pbtype], *map(float, sysm.box),
                                    self.n, float(sysm.temp), cut, sysm.ctoff, sysm.cton, sysm.dcut)

    def __del__(self):
        try:
            self.L.refd_destroy(self.h)
        except Exception:
            pass

    def getE(self):
        e = np.zeros(10)
        self.L.refd_energy(self.h, _p(e, C.c_double))
        return e

    def coords(self):
        x, y, z = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        self.L.refd_get_coords(self.h, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double))
        return x, y, z

    def set_coords(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        self.L.refd_set_coords(self.h, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double))

    def update_list(self):
        self.L.refd_update_list(self.h)

    def list_rows(self):
        cnt = np.zeros(self.n, np.int32)
        tot = self.L.refd_get_list(self.h, _p(cnt, C.c_int), None, 0)
        flat = np.zeros(max(tot, 1), np.int32)
        self.L.refd_get_list(self.h, _p(cnt, C.c_int), _p(flat, C.c_int), tot)
        off = np.concatenate([[0], np.cumsum(cnt)])
        return cnt, [flat[off[i]:off[i + 1]] for i in range(self.n)]

    def excl_rows(self):
        ep = np.zeros(self.n, np.int32)
        n14 = C.c_int(0)
        tot = self.L.refd_get_excl(self.h, _p(ep, C.c_int), None, 0, C.byref(n14), None, 0)
        flat = np.zeros(max(tot, 1), np.int32)
        l14 = np.zeros(max(2 * n14.value, 1), np.int32)
        self.L.refd_get_excl(self.h, _p(ep, C.c_int), _p(flat, C.c_int), tot, C.byref(n14), _p(l14, C.c_int),
                             2 * n14.value)
        off = np.concatenate([[0], np.cumsum(ep)])
        return [flat[off[i]:off[i + 1]] for i in range(self.n)], l14[:2 * n14.value].reshape(-1, 2)

    def mc_run(self, moves, nsteps, seed):
        nt = len(moves)
        typ = np.array([m.type for m in moves], np.int32)
        modes = (C.c_char_p * nt)(*[{1: b"residue", 2: b"all", 4: b"atom"}[m.mode] for m in moves])
        dm = np.array([m.dmax for m in moves]); tg = np.array([m.target for m in moves])
        ea = np.array([m.each for m in moves], np.int32)
        out = np.zeros(nt)
        self.L.refd_mc(self.h, nt, _p(typ, C.c_int), modes, _p(dm, C.c_double), _p(tg, C.c_double),
                       _p(ea, C.c_int), int(nsteps), C.c_uint64(seed), _p(out, C.c_double))
        text = self.L.refd_output().decode()
        trials, accepted = [], []
        for ln in text.splitlines():
            if ln.startswith("TRIALS"):
                p = ln.split()
                trials.append(int(p[1])); accepted.append(int(p[3]))
        return dict(dmax=out, trials=np.array(trials), accepted=np.array(accepted), text=text)


_refmcx = None


def refmcxlib():
    global _refmcx
    if _refmcx is None:
        L = C.CDLL(REFMCXLIB)
        L.refx_output.restype = C.c_char_p
        L.refx_error.restype = C.c_char_p
        L.refx_create.restype = C.c_void_p
        L.refx_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_double, C.c_char_p, C.c_double, C.c_double, C.c_double,
                                  C.c_int, C.c_int]
        L.refx_destroy.argtypes = [C.c_void_p]
        L.refx_energy.argtypes = [C.c_void_p, _dp]
        L.refx_get_coords.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.refx_mc.argtypes = [C.c_void_p, C.c_int, _ip, C.POINTER(C.c_char_p), _dp, _dp, _ip, C.c_int,
                              C.c_uint64, _dp]
        L.refx_mc.restype = C.c_longlong
        L.refx_energies.argtypes = [C.c_void_p, _dp, C.c_longlong]
        _refmcx = L
    return _refmcx


def refmcx_available():
    return os.path.exists(REFMCXLIB)


class ReferenceLoop(Reference):
    """The reference's MC_metropolis::run over FField_MDBAS (gpu=False) or
    over the drop-in FField_MCX of include/mcx_ffield_adapter.hpp
    (gpu=True), through oracle/_ref/libmcref_mcx.so; every getE() the loop
    makes is recorded."""

    def __init__(self, cor, ff, sysm, gpu, device=0):
        self.L = refmcxlib()
        self.n = sysm.natom
        cut = b"full" if sysm.cutmode == 0 else b"switch"
        self.h = self.L.refx_create(cor.encode(), ff.encode(), self.PBC[sysm.pbtype], *map(float, sysm.box),
                                    self.n, float(sysm.temp), cut, sysm.ctoff, sysm.cton, sysm.dcut,
                                    int(bool(gpu)), int(device))
        if not self.h:
            raise RuntimeError("refx_create: " + self.L.refx_error().decode())

    def __del__(self):
        try:
            if self.h:
                self.L.refx_destroy(self.h)
        except Exception:
            pass

    def getE(self):
        e = np.zeros(10)
        self.L.refx_energy(self.h, _p(e, C.c_double))
        return e

    def coords(self):
        x, y, z = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        self.L.refx_get_coords(self.h, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double))
        return x, y, z

    def mc_run(self, moves, nsteps, seed):
        nt = len(moves)
        typ = np.array([m.type for m in moves], np.int32)
        modes = (C.c_char_p * nt)(*[{1: b"residue", 2: b"all", 4: b"atom"}[m.mode] for m in moves])
        dm = np.array([m.dmax for m in moves]); tg = np.array([m.target for m in moves])
        ea = np.array([m.each for m in moves], np.int32)
        out = np.zeros(nt)
        ne = self.L.refx_mc(self.h, nt, _p(typ, C.c_int), modes, _p(dm, C.c_double), _p(tg, C.c_double),
                            _p(ea, C.c_int), int(nsteps), C.c_uint64(seed), _p(out, C.c_double))
        if ne < 0:
            raise RuntimeError("MC_metropolis::run: " + self.L.refx_error().decode())
        e = np.zeros(ne)
        self.L.refx_energies(self.h, _p(e, C.c_double), ne)
        text = self.L.refx_output().decode()
        trials, accepted = [], []
        for ln in text.splitlines():
            if ln.startswith("TRIALS"):
                p = ln.split()
                trials.append(int(p[1])); accepted.append(int(p[3]))
        return dict(dmax=out, trials=np.array(trials), accepted=np.array(accepted), energies=e, text=text)
