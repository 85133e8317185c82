"""ctypes binding of libmcx.so (include/mcx.h) and a host-side mirror of the
reference's energy plugin interface:

  * FField       -- getE(), getE(ener[10]) as getE10(), askListUpdate(st)
                    (include/FField.hpp:95-98)
  * List_nonBonded accessors -- getNeighPair/getNeighList/getExclList/...
                    (include/List_nonBonded.hpp:50-59)
  * MC_metropolis -- run() on the device (src/MC_metropolis.cpp:45-204)

The HIP library is the only compute path: if libmcx.so is missing or no GPU
is visible, construction raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from . import systems as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmcx.so")

MCX_OK = 0
ERRORS = {1: "MCX_ERR_ARG", 2: "MCX_ERR_HIP", 3: "MCX_ERR_LIST_OVERFLOW", 4: "MCX_ERR_BOX_TOO_SMALL",
          5: "MCX_ERR_STATE", 6: "MCX_ERR_UNSUPPORTED"}
MC_REPLAY, MC_PHILOX = 1, 0
LIST_AUTO = -1      # list_every: displacement-triggered rebuilds (include/mcx.h MCX_LIST_AUTO)
MAX_MOVE_TYPES = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MCXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


class MoveTypeC(C.Structure):
    _fields_ = [("type", C.c_int), ("mode", C.c_int), ("dmax", C.c_double), ("target", C.c_double),
                ("each", C.c_int), ("sele", _ip)]


class MCStatsC(C.Structure):
    _fields_ = [("etot", C.c_double), ("trials", C.c_longlong * 8), ("accepted", C.c_longlong * 8),
                ("dmax", C.c_double * 8), ("list_updates", C.c_longlong), ("seconds", C.c_double),
                ("chain_kernel_ms", C.c_double), ("chain_launches", C.c_longlong),
                ("list_kernel_ms", C.c_double), ("list_launches", C.c_longlong)]


class BatchStatsC(C.Structure):
    _fields_ = [("moves", C.c_longlong), ("accepted", C.c_longlong), ("exchange_attempts", C.c_longlong),
                ("exchange_accepted", C.c_longlong), ("sweeps", C.c_longlong), ("seconds", C.c_double),
                ("pair_evals", C.c_longlong)]


_llp = C.POINTER(C.c_longlong)


FRAME_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, _dp, _dp, _dp)


class MCTraceC(C.Structure):
    _fields_ = [("acc", C.POINTER(C.c_ubyte)), ("de", _dp), ("mvtyp", _ip), ("mvinst", _ip), ("disp", _dp),
                ("lnq", _dp), ("save_each", C.c_int), ("on_frame", FRAME_FN), ("user", C.c_void_p)]

# every symbol include/mcx.h declares, with its argument types
SIGNATURES = {
    "mcx_create": ([C.c_int, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "mcx_destroy": ([C.c_void_p], None),
    "mcx_last_error": ([C.c_void_p], C.c_char_p),
    "mcx_set_box": ([C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double], C.c_int),
    "mcx_set_cut": ([C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
    "mcx_set_atoms": ([C.c_void_p] + [_dp] * 6 + [_ip, _ip], C.c_int),
    "mcx_set_bonds": ([C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp, _dp], C.c_int),
    "mcx_set_ub": ([C.c_void_p, C.c_int, _ip, _ip, _dp, _dp], C.c_int),
    "mcx_set_angles": ([C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp], C.c_int),
    "mcx_set_dihedrals": ([C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp], C.c_int),
    "mcx_set_impropers": ([C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp], C.c_int),
    "mcx_build_exclusions": ([C.c_void_p], C.c_int),
    "mcx_get_exclusions": ([C.c_void_p, _ip, _ip, C.c_longlong, C.POINTER(C.c_longlong), _ip, C.c_int, _ip],
                           C.c_int),
    "mcx_set_coords": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "mcx_get_coords": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "mcx_update_list": ([C.c_void_p, C.POINTER(C.c_longlong)], C.c_int),
    "mcx_get_list": ([C.c_void_p, _ip, _ip, C.c_longlong], C.c_int),
    "mcx_list_capacity": ([C.c_void_p, C.c_int, _ip], C.c_int),
    "mcx_energy": ([C.c_void_p, _dp], C.c_int),
    "mcx_energy_noalloc": ([C.c_void_p, _dp], C.c_int),
    "mcx_delta_energy": ([C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
    "mcx_mc_run": ([C.c_void_p, C.POINTER(MoveTypeC), C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int,
                    C.POINTER(MCStatsC), C.POINTER(C.c_ubyte), _dp, _ip, _ip], C.c_int),
    "mcx_mc_run_traced": ([C.c_void_p, C.POINTER(MoveTypeC), C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int,
                           C.POINTER(MCStatsC), C.POINTER(MCTraceC)], C.c_int),
    "mcx_nonbonded_grad": ([C.c_void_p] + [_dp] * 6, C.c_int),
    "mcx_nonbonded_grad_types": ([C.c_void_p, C.c_int, _ip] + [_dp] * 6, C.c_int),
    "mcx_nonbonded_grad_batch": ([C.POINTER(C.c_void_p), C.c_int, C.POINTER(_dp)], C.c_int),
    "mcx_device_coords": ([C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "mcx_coords_updated": ([C.c_void_p], C.c_int),
    "mcx_set_coords_moved": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "mcx_nonbonded_grad_device": ([C.c_void_p, C.c_void_p, C.c_int, _ip, C.c_void_p], C.c_int),
    "mcx_time_kernel": ([C.c_void_p, C.c_int, C.c_int, _dp], C.c_int),
    "mcx_set_rng_stream": ([C.c_void_p, C.c_uint32], C.c_int),
    "mcx_batch_create": ([C.POINTER(C.c_void_p), C.c_int, C.POINTER(MoveTypeC), _ip, C.POINTER(C.c_void_p)],
                         C.c_int),
    "mcx_batch_destroy": ([C.c_void_p], None),
    "mcx_batch_last_error": ([C.c_void_p], C.c_char_p),
    "mcx_batch_reset": ([C.c_void_p, C.c_uint64, C.c_int, C.POINTER(C.c_uint64)], C.c_int),
    "mcx_batch_set_ladder": ([C.c_void_p, _dp], C.c_int),
    "mcx_batch_set_sweep": ([C.c_void_p, _ip], C.c_int),
    "mcx_batch_run": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(BatchStatsC)], C.c_int),
    "mcx_batch_moves": ([C.c_void_p, C.c_int, C.POINTER(C.c_ubyte), _dp], C.c_int),
    "mcx_batch_state": ([C.c_void_p, _dp, _ip, _llp, _llp, _dp], C.c_int),
    "mcx_batch_exchange_stats": ([C.c_void_p, _llp, _llp], C.c_int),
    "mcx_batch_set_config": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "mcx_batch_tune": ([C.c_void_p, _ip, _ip, _dp], C.c_int),
}

_lib = None


def load_library(path: str = LIBPATH):
    """Loads libmcx.so and binds every declared entry point (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("libmcx.so not built (run __graft_entry__.build()); no fallback path exists")
        L = C.CDLL(path)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Handle:
    """One mcx_ctx: a system resident on one GPU."""

    def __init__(self, natom: int, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.mcx_create(int(natom), int(device), C.byref(h))
        if rc != MCX_OK:
            raise MCXError(rc, "mcx_create failed (no usable HIP device %d?)" % device)
        self.h = h
        self.n = natom

    def close(self):
        if getattr(self, "h", None):
            self.L.mcx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != MCX_OK:
            raise MCXError(rc, self.L.mcx_last_error(self.h).decode())

    # ---- set-up ----
    def set_box(self, pbtype, a, b=None, c=None):
        b = a if b is None else b
        c = a if c is None else c
        self._chk(self.L.mcx_set_box(self.h, int(pbtype), float(a), float(b), float(c)))

    def set_cut(self, cutmode, ctoff=12.0, cton=10.0, dcut=2.0, temp=300.0):
        self._chk(self.L.mcx_set_cut(self.h, int(cutmode), float(ctoff), float(cton), float(dcut), float(temp)))

    def set_atoms(self, charge, eps, sig, eps14=None, sig14=None, mass=None, frozen=None, resid=None):
        arrs = [_d(charge), _d(eps), _d(sig), None if eps14 is None else _d(eps14),
                None if sig14 is None else _d(sig14), None if mass is None else _d(mass)]
        fr = None if frozen is None else _i(frozen)
        rs = None if resid is None else _i(resid)
        self._chk(self.L.mcx_set_atoms(self.h, *[_p(a, C.c_double) for a in arrs], _p(fr, C.c_int), _p(rs, C.c_int)))

    def set_terms(self, s: S.System):
        b, u, a, d, m = s.bonds, s.ub, s.angles, s.dihedrals, s.impropers
        keep = []

        def I(x):
            x = _i(x); keep.append(x); return _p(x, C.c_int)

        def D(x):
            x = _d(x); keep.append(x); return _p(x, C.c_double)

        self._chk(self.L.mcx_set_bonds(self.h, len(b["a"]), I(b["a"]), I(b["b"]), I(b["type"]), D(b["k"]),
                                       D(b["r0"]), D(b["beta"])))
        self._chk(self.L.mcx_set_ub(self.h, len(u["a"]), I(u["a"]), I(u["b"]), D(u["k"]), D(u["r0"])))
        self._chk(self.L.mcx_set_angles(self.h, len(a["a"]), I(a["a"]), I(a["b"]), I(a["c"]), D(a["k"]),
                                        D(a["t0"])))
        self._chk(self.L.mcx_set_dihedrals(self.h, len(d["a"]), I(d["a"]), I(d["b"]), I(d["c"]), I(d["d"]),
                                           I(d["type"]), D(d["k"]), D(d["phi0"]), D(d["mult"])))
        self._chk(self.L.mcx_set_impropers(self.h, len(m["a"]), I(m["a"]), I(m["b"]), I(m["c"]), I(m["d"]),
                                           I(m["type"]), D(m["k"]), D(m["phi0"]), D(m["mult"])))

    def build_exclusions(self):
        self._chk(self.L.mcx_build_exclusions(self.h))

    def set_coords(self, x, y, z):
        x, y, z = _d(x), _d(y), _d(z)
        self._chk(self.L.mcx_set_coords(self.h, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double)))

    def set_coords_moved(self, x, y, z):
        """Coordinates changed by moves: the pair list stays as built (the
        reference's getE between list updates)."""
        x, y, z = _d(x), _d(y), _d(z)
        self._chk(self.L.mcx_set_coords_moved(self.h, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double)))

    def get_coords(self):
        x, y, z = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        self._chk(self.L.mcx_get_coords(self.h, _p(x, C.c_double), _p(y, C.c_double), _p(z, C.c_double)))
        return x, y, z

    # ---- lists ----
    def update_list(self) -> int:
        n = C.c_longlong(0)
        self._chk(self.L.mcx_update_list(self.h, C.byref(n)))
        return n.value

    def get_list(self):
        cnt = np.zeros(self.n, np.int32)
        self._chk(self.L.mcx_get_list(self.h, _p(cnt, C.c_int), None, 0))
        flat = np.zeros(max(int(cnt.sum()), 1), np.int32)
        self._chk(self.L.mcx_get_list(self.h, _p(cnt, C.c_int), _p(flat, C.c_int), flat.size))
        return cnt, flat[: int(cnt.sum())]

    def list_capacity(self, set_to: int = 0) -> int:
        """Row capacity of the pair list; set_to > 0 makes it exactly that
        (List_nonBonded's sizeList) and marks the list stale."""
        cur = C.c_int(0)
        self._chk(self.L.mcx_list_capacity(self.h, int(set_to), C.byref(cur)))
        return cur.value

    def get_exclusions(self):
        ep = np.zeros(self.n, np.int32)
        nidx = C.c_longlong(0)
        n14 = C.c_int(0)
        self._chk(self.L.mcx_get_exclusions(self.h, _p(ep, C.c_int), None, 0, C.byref(nidx), None, 0, C.byref(n14)))
        idx = np.zeros(max(nidx.value, 1), np.int32)
        p14 = np.zeros(max(2 * n14.value, 1), np.int32)
        self._chk(self.L.mcx_get_exclusions(self.h, _p(ep, C.c_int), _p(idx, C.c_int), idx.size, C.byref(nidx),
                                            _p(p14, C.c_int), p14.size, C.byref(n14)))
        return ep, idx[: nidx.value], p14[: 2 * n14.value].reshape(-1, 2)

    # ---- energies ----
    def energy(self):
        e = np.zeros(10)
        self._chk(self.L.mcx_energy(self.h, _p(e, C.c_double)))
        return e

    def energy_noalloc(self):
        """mcx_energy_noalloc: the energy with no device allocation (a
        current pair list is required in switched mode)."""
        e = np.zeros(10)
        self._chk(self.L.mcx_energy_noalloc(self.h, _p(e, C.c_double)))
        return e

    def delta_energy(self, first, last, nx, ny, nz):
        nx, ny, nz = _d(nx), _d(ny), _d(nz)
        out = C.c_double(0)
        self._chk(self.L.mcx_delta_energy(self.h, int(first), int(last), _p(nx, C.c_double), _p(ny, C.c_double),
                                          _p(nz, C.c_double), C.byref(out)))
        return out.value

    def grad(self):
        out = [np.zeros(self.n) for _ in range(6)]
        self._chk(self.L.mcx_nonbonded_grad(self.h, *[_p(a, C.c_double) for a in out]))
        return out

    def grad_types(self, type_of_atom, ntypes=None):
        """Forces per atom and parameter gradients summed per type."""
        ty = _i(type_of_atom)
        nt = int(ty.max()) + 1 if ntypes is None else int(ntypes)
        gx, gy, gz = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        ge, gs, gq = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        self._chk(self.L.mcx_nonbonded_grad_types(self.h, nt, _p(ty, C.c_int), *[_p(a, C.c_double) for a in
                                                                                (gx, gy, gz, ge, gs, gq)]))
        return gx, gy, gz, ge, gs, gq

    def device_coords(self, which) -> int:
        """Device address of the master coordinate array x (0), y (1), z (2)."""
        p = C.c_void_p(0)
        self._chk(self.L.mcx_device_coords(self.h, int(which), C.byref(p)))
        return int(p.value)

    def coords_updated(self):
        """The device coordinate arrays were written in place: refresh the
        fixed-point mirror on the device, mark the list stale."""
        self._chk(self.L.mcx_coords_updated(self.h))

    def grad_device(self, gpos_ptr, type_of_atom=None, ntypes=0, gtype_ptr=0):
        """Adjoint into device memory: gpos_ptr -> 3 n doubles (dE/dx, dE/dy,
        dE/dz); with types, gtype_ptr -> 5 ntypes doubles per type (eps,
        sig, q, eps14, sig14)."""
        ty = _i(type_of_atom) if type_of_atom is not None else None
        self._chk(self.L.mcx_nonbonded_grad_device(self.h, C.c_void_p(int(gpos_ptr)), int(ntypes),
                                                   _p(ty, C.c_int) if ty is not None else None,
                                                   C.c_void_p(int(gtype_ptr)) if gtype_ptr else None))

    def time_kernel(self, which, reps=20):
        ms = C.c_double(0)
        self._chk(self.L.mcx_time_kernel(self.h, int(which), int(reps), C.byref(ms)))
        return ms.value

    def set_rng_stream(self, stream: int):
        self._chk(self.L.mcx_set_rng_stream(self.h, C.c_uint32(stream)))

    def mc_run(self, moves: Sequence[S.MoveType], nsteps, seed, rng_mode=MC_REPLAY, list_every=50, record=True,
               save_each=0, on_frame=None):
        """Device MC_metropolis::run.  on_frame(step, etot, x, y, z) is called
        every save_each steps (numpy copies of the coordinates)."""
        nt = len(moves)
        arr, keep = _move_array(moves)
        st = MCStatsC()
        acc = np.zeros(max(nsteps, 1), np.uint8) if record else None
        de = np.zeros(max(nsteps, 1)) if record else None
        ty = np.zeros(max(nsteps, 1), np.int32) if record else None
        ins = np.zeros(max(nsteps, 1), np.int32) if record else None
        disp = np.zeros((max(nsteps, 1), 3)) if record else None
        lnq = np.zeros(max(nsteps, 1)) if record else None
        n = self.n
        cb = FRAME_FN()
        if on_frame is not None and save_each > 0:
            def _frame(user, step, etot, x, y, z):
                on_frame(int(step), float(etot), *(np.ctypeslib.as_array(a, (n,)).copy() for a in (x, y, z)))
            cb = FRAME_FN(_frame)
        tr = MCTraceC(_p(acc, C.c_ubyte), _p(de, C.c_double), _p(ty, C.c_int), _p(ins, C.c_int),
                      _p(disp, C.c_double), _p(lnq, C.c_double), int(save_each), cb, None)
        self._chk(self.L.mcx_mc_run_traced(self.h, arr, nt, int(nsteps), C.c_uint64(seed), int(rng_mode),
                                           int(list_every), C.byref(st), C.byref(tr)))
        out = dict(etot=st.etot, trials=np.array(st.trials[:nt]), accepted=np.array(st.accepted[:nt]),
                   dmax=np.array(st.dmax[:nt]), list_updates=st.list_updates, seconds=st.seconds,
                   chain_kernel_ms=st.chain_kernel_ms, chain_launches=st.chain_launches,
                   list_kernel_ms=st.list_kernel_ms, list_launches=st.list_launches)
        if record:
            out.update(acc=acc[:nsteps], de=de[:nsteps], mvtyp=ty[:nsteps], mvinst=ins[:nsteps],
                       disp=disp[:nsteps], lnq=lnq[:nsteps])
        return out


def _move_array(moves: Sequence[S.MoveType]):
    arr = (MoveTypeC * len(moves))()
    keep = []
    for k, m in enumerate(moves):
        arr[k].type, arr[k].mode, arr[k].dmax, arr[k].target, arr[k].each = m.type, m.mode, m.dmax, m.target, m.each
        if m.sele is not None:
            sa = _i(m.sele); keep.append(sa); arr[k].sele = _p(sa, C.c_int)
    return arr, keep


def grad_batch(handles: Sequence[Handle]):
    """Gradients of many systems in one launch: a list of (gx, gy, gz, geps,
    gsig, gq) per system."""
    L = load_library()
    hs = (C.c_void_p * len(handles))(*[h.h for h in handles])
    outs = [np.zeros(6 * h.n) for h in handles]
    ptrs = (_dp * len(handles))(*[_p(o, C.c_double) for o in outs])
    rc = L.mcx_nonbonded_grad_batch(hs, len(handles), ptrs)
    if rc != MCX_OK:
        raise MCXError(rc, L.mcx_last_error(handles[0].h).decode())
    return [tuple(o.reshape(6, h.n)) for o, h in zip(outs, handles)]


def from_system(s: S.System, device: int = 0) -> Handle:
    """Creates a handle holding System `s` (the Parser_XML + IO_MDBAS +
    List_nonBonded set-up sequence, src/Parser.cpp:77-215)."""
    h = Handle(s.natom, device)
    h.set_box(s.pbtype, *s.box)
    h.set_cut(s.cutmode, s.ctoff, s.cton, s.dcut, s.temp)
    h.set_atoms(s.charge, s.eps, s.sig, s.eps14, s.sig14, s.mass, s.frozen, s.resid)
    h.set_terms(s)
    h.build_exclusions()
    h.set_coords(s.x, s.y, s.z)
    return h


class FField:
    """Mirror of the reference's FField call surface (include/FField.hpp:52-153)
    backed by the gfx950 kernels."""

    def __init__(self, s: S.System, device: int = 0):
        self.sys = s
        self.h = from_system(s, device)
        if s.cutmode != S.CUT_FULL:
            self.h.update_list()      # List_nonBonded ctor builds the list (src/List_nonBonded.cpp:62-100)

    def getE(self) -> float:
        return float(self.h.energy()[0])

    def getE10(self) -> np.ndarray:
        """getE(double ener[10]): {tot, pot, kin, elec, vdw, bond, ang, ub, dihe, impr}."""
        return self.h.energy()

    def askListUpdate(self, st: int = 0):
        """FField::askListUpdate (src/FField.cpp:222-235): no-op in FULL mode."""
        if self.sys.cutmode != S.CUT_FULL:
            self.h.update_list()

    def getCutMode(self):
        return self.sys.cutmode

    # List_nonBonded accessors (src/List_nonBonded.cpp:1343-1376)
    def getNeighPair(self):
        return self.h.get_list()[0]

    def getNeighList(self):
        cnt, flat = self.h.get_list()
        off = np.concatenate([[0], np.cumsum(cnt)])
        return [flat[off[i]:off[i + 1]] for i in range(self.sys.natom)]

    def getNeighOrder(self):
        return np.arange(self.sys.natom, dtype=np.int32)

    def getExclPair(self):
        return self.h.get_exclusions()[0]

    def getExclList(self):
        ep, idx, _ = self.h.get_exclusions()
        off = np.concatenate([[0], np.cumsum(ep)])
        return [idx[off[i]:off[i + 1]] for i in range(self.sys.natom)]

    def getNeighList14(self):
        return self.h.get_exclusions()[2].reshape(-1)

    def getNPair14(self):
        return len(self.h.get_exclusions()[2])

    def set_coords(self, x, y, z):
        self.h.set_coords(x, y, z)

    def get_coords(self):
        return self.h.get_coords()


class MC_metropolis:
    """MC_metropolis (src/MC_metropolis.cpp:29-204) driving FField on the device."""

    def __init__(self, ff: FField, moves: Optional[Sequence[S.MoveType]] = None, nsteps: int = 0, seed: int = 1,
                 rng_mode: int = MC_REPLAY, list_every: int = 50):
        self.ff, self.moves = ff, list(moves if moves is not None else ff.sys.moves)
        self.nsteps, self.seed, self.rng_mode, self.list_every = nsteps, seed, rng_mode, list_every

    def run(self, record=True):
        if self.nsteps == 0:
            return None       # "Dummy MC simulation because nsteps = 0" (src/MC.cpp:177-181)
        return self.ff.h.mc_run(self.moves, self.nsteps, self.seed, self.rng_mode, self.list_every, record)


class Batch:
    """Replicas run together: one Metropolis chain per member handle, one
    persistent kernel launch per sweep, replica exchange on the device
    (mcx_batch_* of include/mcx.h).  The handles must stay alive as long as
    the batch; the batch keeps references to them."""

    def __init__(self, handles: Sequence[Handle], moves):
        """moves: one list of MoveType shared by every replica (selections
        None), or one list per replica."""
        self.L = load_library()
        self.handles = list(handles)
        self.nrep = len(self.handles)
        per = [list(m) for m in moves] if len(moves) and isinstance(moves[0], (list, tuple)) \
            else [list(moves)] * self.nrep
        assert len(per) == self.nrep
        self.ntypes = np.array([len(m) for m in per], dtype=np.int32)
        hs = (C.c_void_p * self.nrep)(*[h.h for h in self.handles])
        arr, keep = _move_array([m for ms in per for m in ms])
        b = C.c_void_p()
        rc = self.L.mcx_batch_create(hs, self.nrep, arr, _p(self.ntypes, C.c_int), C.byref(b))
        if rc != MCX_OK:
            raise MCXError(rc, self.L.mcx_last_error(self.handles[0].h).decode())
        self.b = b

    def close(self):
        if getattr(self, "b", None):
            self.L.mcx_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != MCX_OK:
            raise MCXError(rc, self.L.mcx_batch_last_error(self.b).decode())

    def reset(self, seed=0, rng_mode=MC_PHILOX, seeds=None):
        sp = None
        if seeds is not None:
            sa = np.ascontiguousarray(seeds, dtype=np.uint64)
            sp = sa.ctypes.data_as(C.POINTER(C.c_uint64))
        self._chk(self.L.mcx_batch_reset(self.b, C.c_uint64(seed), int(rng_mode), sp))

    def set_ladder(self, temps):
        t = _d(temps)
        assert t.size == self.nrep
        self._chk(self.L.mcx_batch_set_ladder(self.b, _p(t, C.c_double)))

    def set_sweep(self, moves_per_sweep):
        m = _i(moves_per_sweep)
        assert m.size == self.nrep
        self._chk(self.L.mcx_batch_set_sweep(self.b, _p(m, C.c_int)))

    def run(self, nsweeps, exchange_every=0, use_graph=True):
        st = BatchStatsC()
        self._chk(self.L.mcx_batch_run(self.b, int(nsweeps), int(exchange_every), int(bool(use_graph)), C.byref(st)))
        return dict(moves=st.moves, accepted=st.accepted, exchange_attempts=st.exchange_attempts,
                    exchange_accepted=st.exchange_accepted, sweeps=st.sweeps, seconds=st.seconds,
                    pair_evals=st.pair_evals)

    def moves(self, nmoves, record=True):
        acc = np.zeros((self.nrep, nmoves), np.uint8) if record else None
        de = np.zeros((self.nrep, nmoves)) if record else None
        self._chk(self.L.mcx_batch_moves(self.b, int(nmoves), _p(acc, C.c_ubyte), _p(de, C.c_double)))
        return acc, de

    def state(self):
        """trials / accepted / dmax: (nrep, MAX_MOVE_TYPES), zero past each
        replica's own move types."""
        R, T = self.nrep, MAX_MOVE_TYPES
        etot = np.zeros(R)
        slot = np.zeros(R, np.int32)
        tr = np.zeros((R, T), np.int64)
        ac = np.zeros((R, T), np.int64)
        dm = np.zeros((R, T))
        self._chk(self.L.mcx_batch_state(self.b, _p(etot, C.c_double), _p(slot, C.c_int), _p(tr, C.c_longlong),
                                         _p(ac, C.c_longlong), _p(dm, C.c_double)))
        return dict(etot=etot, slot=slot, trials=tr, accepted=ac, dmax=dm)

    def exchange_stats(self):
        n = max(self.nrep - 1, 1)
        a = np.zeros(n, np.int64)
        c = np.zeros(n, np.int64)
        self._chk(self.L.mcx_batch_exchange_stats(self.b, _p(a, C.c_longlong), _p(c, C.c_longlong)))
        return a[: self.nrep - 1], c[: self.nrep - 1]

    def set_config(self, threads, unroll):
        self._chk(self.L.mcx_batch_set_config(self.b, int(threads), int(unroll)))

    def tune(self):
        t, u, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        self._chk(self.L.mcx_batch_tune(self.b, C.byref(t), C.byref(u), C.byref(ms)))
        return t.value, u.value, ms.value
