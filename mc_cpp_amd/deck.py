"""The reference's XML input decks (Parser_XML, src/Parser.cpp:29-279) read
into a System, its move types and the Metropolis run parameters, and run on
the device with the reference's trajectory and energy files.

Semantics kept from the reference:
  * root <inputFile title=...>; its sections are handled in document order
    (<pbc>, <ensemble>, <forcefield>, <atomlist>, <movelist>, <mc>,
    <benchmark>); an unknown section is ignored with a warning (:272-275);
  * node and attribute names are case sensitive, values are blank-stripped
    and lower-cased where the reference does so (pbc_type, ens_type,
    ff_type, cut_type, list_mode, move_type, move_mode, sel_mode), paths and
    selection strings are blank-stripped only;
  * a missing attribute keeps the parser's current value: the defaults of
    include/Parser.hpp:115-148 at first, then whatever an earlier node set
    (so a <move> without dmax_value inherits the previous move's);
  * pbc_type cubic uses a for all three edges (src/PerConditions.cpp:101-
    110); an unknown pbc_type means no PBC;
  * <atomlist list_mode="manual"> creates N identical atoms with the given
    charge, lj_epsilon and lj_sigma as AtomList values (no unit
    conversion) at the origin (:166-177); "file" reads the MDBAS COR and
    forfield files (IO_MDBAS);
  * selections (src/Selection.cpp:27-152): all, none, residue_id (COR
    column 2 equal to the leading integer of sele), residue_name (COR
    column 3 equal to sele), atom_idx (COR column 1);
  * move types (src/List_Moves.cpp:170-211, 257-300): trn / rot, modes
    residue / all / atom; an unknown type or mode is skipped with a
    warning; rot with atom mode is an error;
  * <mc nsteps save_each seed>: save_each is read only when nsteps != 0;
    seed 0 asks for a random seed (src/MC.cpp:110-117).
Paths in the deck are resolved against the deck's directory (the reference
resolves them against its working directory, which is the decks' own
directory in its test layout).

Trajectory and energies (src/MC.cpp:93-108, src/MC_metropolis.cpp:157-163):
every save_each steps "tr.xyz" receives "<natom>", "#At step <st>" and one
"<symbol>\\t%10.5f\\t%10.5f\\t%10.5f" line per atom, and "ener.dat" one
"<st>\\t%f" line with the running total energy (iu).
"""
from __future__ import annotations

import dataclasses
import os
import re
import xml.etree.ElementTree as ET
from typing import List, Optional

import numpy as np

from . import systems as S


class DeckError(ValueError):
    """An input the reference rejects with exit()."""


def _blank(v: str) -> str:
    return re.sub(r"\s+", "", v)


def _leading_int(v: str) -> Optional[int]:
    m = re.match(r"[+-]?\d+", _blank(v))
    return int(m.group(0)) if m else None


@dataclasses.dataclass
class Deck:
    title: str
    system: S.System
    nsteps: int
    save_each: int
    seed: int
    benchmark: bool
    path: str
    warnings: List[str]


def read_deck(path: str, base_dir: Optional[str] = None) -> Deck:
    base = base_dir if base_dir is not None else os.path.dirname(os.path.abspath(path))
    root = ET.parse(path).getroot()
    if root.tag != "inputFile":
        raise DeckError("root node must be <inputFile>")
    title = root.attrib.get("title", "")
    # parser state with the defaults of include/Parser.hpp:115-148
    st = dict(pbc_type="none", a=0.0, b=0.0, c=0.0, ens_type="nvt", N=0, T=0.0, ff_type="", ff_file="",
              cut_type="full", cutoff=12.0, cuton=10.0, delta_cut=2.0, list_mode="file", symbol="",
              charge=0.0, lj_epsilon=1.0, lj_sigma=1.0, cor_path="", move_type="trn", move_mode="all",
              sel_mode="all", sele="", dmax_value=0.5, dmax_target=50.0, dmax_each=100, nsteps=10000,
              save_each=100, seed=0)
    conv = dict(a=float, b=float, c=float, N=int, T=float, cutoff=float, cuton=float, delta_cut=float,
                charge=float, lj_epsilon=float, lj_sigma=float, dmax_value=float, dmax_target=float,
                dmax_each=int, nsteps=int, save_each=int, seed=int)
    warnings = []

    def take(node, name):
        if name in node.attrib:
            v = node.attrib[name]
            st[name] = conv[name](_blank(v)) if name in conv else v

    pbtype = S.PBC_NONE
    box = np.zeros(3)
    system = None
    moves = []
    benchmark = False
    cutmode = S.CUT_FULL
    for node in root:
        tag = node.tag
        if not isinstance(tag, str):
            continue                                   # comments
        if tag == "pbc":
            take(node, "pbc_type")
            kind = _blank(st["pbc_type"]).lower()
            if kind != "none":
                for k in ("a", "b", "c"):
                    take(node, k)
            if kind == "cubic":
                pbtype, box = S.PBC_CUBIC, np.array([st["a"]] * 3)
            elif kind == "orbic":
                pbtype, box = S.PBC_ORBIC, np.array([st["a"], st["b"], st["c"]])
            else:
                if kind != "none":
                    warnings.append("%s is not a valid PBC type; no PBC" % kind)
                pbtype, box = S.PBC_NONE, np.array([st["a"], st["b"], st["c"]])
        elif tag == "ensemble":
            for k in ("ens_type", "N", "T"):
                take(node, k)
            if _blank(st["ens_type"]).lower() != "nvt":
                raise DeckError("only the NVT ensemble is supported")
        elif tag == "forcefield":
            for k in ("ff_type", "ff_file", "cut_type"):
                take(node, k)
            cut = _blank(st["cut_type"]).lower()
            if cut == "switch":
                for k in ("cutoff", "cuton", "delta_cut"):
                    take(node, k)
                cutmode = S.CUT_SWITCH
            elif cut == "full":
                cutmode = S.CUT_FULL
            else:
                raise DeckError("cut_type must be 'switch' or 'full'")
        elif tag == "atomlist":
            take(node, "list_mode")
            mode = _blank(st["list_mode"]).lower()
            common = dict(pbtype=pbtype, cutmode=cutmode, ctoff=st["cutoff"], cton=st["cuton"],
                          dcut=st["delta_cut"], temp=st["T"])
            if mode == "manual":
                for k in ("symbol", "charge", "lj_epsilon", "lj_sigma"):
                    take(node, k)
                n = st["N"]
                system = S.System(natom=n, box=box.copy(), charge=np.full(n, st["charge"]),
                                  eps=np.full(n, st["lj_epsilon"]), sig=np.full(n, st["lj_sigma"]), **common)
                system.symbols = [st["symbol"]] * n
                system.resnames = [""] * n
            elif mode == "file":
                take(node, "cor_path")
                if _blank(st["ff_type"]).lower() != "mdbas":
                    raise DeckError("only MDBAS forcefield and coordinates (COR) are supported")
                cor = os.path.join(base, _blank(st["cor_path"]))
                ff = os.path.join(base, _blank(st["ff_file"]))
                system = S.read_mdbas(cor, ff, box.copy(), **common)
                if system.natom != st["N"]:
                    raise DeckError("number of atoms in %s (%d) differs from the ensemble's N (%d)"
                                    % (cor, system.natom, st["N"]))
            else:
                raise DeckError("atomlist list_mode must be 'manual' or 'file'")
        elif tag == "movelist":
            if system is None:
                raise DeckError("<movelist> before <atomlist>")
            for mv in node.findall("move"):
                for k in ("move_type", "move_mode", "sel_mode", "dmax_value", "dmax_target", "dmax_each"):
                    take(mv, k)
                smode = _blank(st["sel_mode"]).lower()
                if smode not in ("none", "all"):
                    take(mv, "sele")
                mtyp = _blank(st["move_type"]).lower()
                mmode = _blank(st["move_mode"]).lower()
                if mtyp not in ("trn", "rot"):
                    warnings.append("%s is not a valid type of move; skipped" % mtyp)
                    continue
                sele = selection(system, smode, st["sele"])
                if mmode not in ("residue", "all", "atom"):
                    warnings.append("move mode %s is not available; skipped" % mmode)
                    continue
                if mmode == "atom" and mtyp == "rot":
                    raise DeckError("rotation with single atom mode is useless")
                moves.append(S.MoveType(S.MV_TRN if mtyp == "trn" else S.MV_ROT,
                                        {"residue": S.MODE_RESIDUE, "all": S.MODE_ALL, "atom": S.MODE_ATOM}[mmode],
                                        st["dmax_value"], st["dmax_target"], st["dmax_each"], sele))
        elif tag == "mc":
            take(node, "nsteps")
            if st["nsteps"] != 0:
                take(node, "save_each")
            take(node, "seed")
        elif tag == "benchmark":
            benchmark = True
        else:
            warnings.append('unknown section "<%s>" ignored' % tag)
    if system is None:
        raise DeckError("no <atomlist> in the deck")
    system.moves = moves
    return Deck(title=title, system=system, nsteps=st["nsteps"], save_each=st["save_each"], seed=st["seed"],
                benchmark=benchmark, path=os.path.abspath(path), warnings=warnings)


def selection(s: S.System, mode: str, sele: str) -> np.ndarray:
    """0/1 per atom as Selection::select_main builds it (src/Selection.cpp:73-152)."""
    n = s.natom
    if mode == "all":
        return np.ones(n, dtype=np.int32)
    if mode == "none":
        return np.zeros(n, dtype=np.int32)
    if mode == "residue_id":
        num = _leading_int(sele)
        return (s.resid == num).astype(np.int32) if num is not None else np.zeros(n, dtype=np.int32)
    if mode == "residue_name":
        name = _blank(sele)
        return np.array([1 if (r == name) else 0 for r in (s.resnames or [""] * n)], dtype=np.int32)
    if mode == "atom_idx":
        num = _leading_int(sele)
        return (s.atom_ids == num).astype(np.int32) if num is not None else np.zeros(n, dtype=np.int32)
    raise DeckError("the following 'sel_mode' is unknown : %s" % mode)


def run_deck(deck: Deck, out_dir: Optional[str] = None, rng_mode=None, list_every: int = 50, device: int = 0):
    """MC_metropolis::run of the deck on the device (src/MC_metropolis.cpp:45-173).
    With save_each != 0 writes tr.xyz and ener.dat into out_dir (default:
    the working directory).  Returns the run statistics and the seed used."""
    from . import mcx
    s = deck.system
    ff = mcx.FField(s, device=device)
    seed = deck.seed if deck.seed != 0 else int.from_bytes(os.urandom(4), "little")
    out = {"seed": seed}
    if deck.benchmark:
        import time
        t0 = time.perf_counter()
        out["benchmark_energy"] = ff.getE()
        out["benchmark_seconds"] = time.perf_counter() - t0
    if deck.nsteps == 0:
        return out
    rng_mode = mcx.MC_REPLAY if rng_mode is None else rng_mode
    out_dir = out_dir or os.getcwd()
    files = None
    if deck.save_each != 0:
        files = (open(os.path.join(out_dir, "tr.xyz"), "w"), open(os.path.join(out_dir, "ener.dat"), "w"))
    sym = s.symbols or ["X"] * s.natom

    def frame(step, etot, x, y, z):
        xyz, ene = files
        xyz.write("%d\n#At step %d\n" % (s.natom, step))
        xyz.write("".join("%s\t%10.5f\t%10.5f\t%10.5f\n" % (sym[i], x[i], y[i], z[i]) for i in range(s.natom)))
        ene.write("%d\t%f\n" % (step, etot))

    try:
        res = ff.h.mc_run(s.moves, deck.nsteps, seed, rng_mode=rng_mode, list_every=list_every, record=False,
                          save_each=deck.save_each if files else 0, on_frame=frame if files else None)
    finally:
        if files:
            for f in files:
                f.close()
    out.update(res)
    return out
