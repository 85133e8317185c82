"""XML input decks (Parser_XML, src/Parser.cpp:29-279) read into systems,
move types and run parameters; the decks' energies against the energies the
reference's own build computed for the same inputs (tests/golden/golden.json),
and a deck run on the device with its tr.xyz / ener.dat frames.

The fixture decks under tests/golden/decks carry the sections and values of
the reference's argon_box and diala_vac decks; the other decks exercise the
parser rules (defaults carried between nodes, selection modes, manual atom
lists, rejected inputs).  The reference's own decks are parsed too when the
reference tree is present.
"""
import json
import os
import re

import numpy as np
import pytest

from mc_cpp_amd import deck as D
from mc_cpp_amd import systems
from oracle.oracle import Oracle, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")
REF_TEST = "/root/reference/test"


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def oracle_lib():
    return lib()


def _check_energy(s, ref_energy):
    o = Oracle(s.copy())
    if s.cutmode == systems.CUT_SWITCH:
        o.update_list()
    e = o.getE()
    for k in range(10):
        assert abs(e[k] - ref_energy[k]) <= 1e-9 * max(abs(ref_energy[k]), 1e-6), (k, e[k], ref_energy[k])


def test_argon_deck(golden, oracle_lib):
    d = D.read_deck(os.path.join(DECKS, "argon_box.xml"))
    s = d.system
    assert d.title == "Argon atoms in a cubic box" and d.benchmark and d.nsteps == 0 and d.seed == 0
    assert s.natom == 256 and s.pbtype == systems.PBC_CUBIC and np.allclose(s.box, 30.0)
    assert s.cutmode == systems.CUT_SWITCH and (s.ctoff, s.cton, s.dcut, s.temp) == (12.0, 10.0, 2.0, 300.0)
    (m,) = s.moves
    assert (m.type, m.mode, m.dmax, m.target, m.each) == (systems.MV_TRN, systems.MODE_ATOM, 0.5, 50.0, 100)
    assert int(m.sele.sum()) == 256
    _check_energy(s, golden["argon_box"]["energy"])


def test_diala_deck(golden, oracle_lib):
    d = D.read_deck(os.path.join(DECKS, "diala_vac.xml"))
    s = d.system
    assert s.natom == 12 and s.cutmode == systems.CUT_FULL
    assert [(m.type, m.mode, m.dmax) for m in s.moves] == [(systems.MV_TRN, systems.MODE_ATOM, 0.15)]
    assert list(s.atom_ids) == list(range(1, 13)) and list(s.resid) == [1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3, 3]
    _check_energy(s, golden["diala_vac"]["energy"])


def _write(tmp_path, body, name="deck.xml"):
    p = tmp_path / name
    p.write_text('<?xml version="1.0"?>\n<inputFile title="t">\n' + body + "\n</inputFile>\n")
    return str(p)


def _diala_head():
    ff = os.path.join(HERE, "golden", "diala_vac", "ff.dat")
    cor = os.path.join(HERE, "golden", "diala_vac", "ala2.cor")
    return ('<pbc pbc_type=" Cubic " a="30.0"/>\n<ensemble ens_type="NVT" N="12" T="250"/>\n'
            '<forcefield ff_type="MDBAS" ff_file="%s" cut_type="Switch" cutoff="9.0"/>\n'
            '<atomlist list_mode="file" cor_path="%s"/>\n' % (ff, cor))


def test_parser_rules(tmp_path):
    body = _diala_head() + '''
<movelist>
  <move move_type="rot" move_mode="residue" sel_mode="residue_name" sele=" ALA " dmax_value="20" dmax_each="10" dmax_target="40"/>
  <move move_type="TRN" move_mode="atom" sel_mode="atom_idx" sele="7"/>
  <move move_type="trn" move_mode="residue" sel_mode="residue_id" sele="3 junk"/>
  <move move_type="tors" move_mode="residue" sel_mode="all"/>
  <move move_type="trn" move_mode="sideways" sel_mode="none"/>
  <move move_type="trn" move_mode="all" sel_mode="none"/>
</movelist>
<analysis/>
<mc nsteps="0" save_each="7" seed="42"/>'''
    d = D.read_deck(_write(tmp_path, body))
    s = d.system
    # cutoff given, cuton and delta_cut keep their defaults; cubic uses a
    assert (s.ctoff, s.cton, s.dcut, s.temp) == (9.0, 10.0, 2.0, 250.0) and np.allclose(s.box, 30.0)
    assert [(m.type, m.mode) for m in s.moves] == [(systems.MV_ROT, systems.MODE_RESIDUE),
                                                    (systems.MV_TRN, systems.MODE_ATOM),
                                                    (systems.MV_TRN, systems.MODE_RESIDUE),
                                                    (systems.MV_TRN, systems.MODE_ALL)]
    # dmax values carried over from the first move
    assert all((m.dmax, m.each, m.target) == (20.0, 10, 40.0) for m in s.moves)
    sel = [list(np.nonzero(m.sele)[0]) for m in s.moves]
    assert sel == [[3, 4, 5, 6, 7, 8], [6], [9, 10, 11], []]
    assert len(d.warnings) == 3        # tors, sideways, <analysis>
    # save_each is read only when nsteps != 0
    assert (d.nsteps, d.save_each, d.seed) == (0, 100, 42)


def test_manual_atoms_and_rejections(tmp_path):
    body = ('<pbc pbc_type="orbic" a="20" b="21" c="22"/>\n<ensemble ens_type="nvt" N="5" T="90"/>\n'
            '<forcefield ff_type="mdbas" cut_type="full"/>\n'
            '<atomlist list_mode="manual" symbol="Ar" charge="0.0" lj_epsilon="0.238" lj_sigma="1.705"/>\n'
            '<mc nsteps="10" save_each="5" seed="3"/>')
    d = D.read_deck(_write(tmp_path, body))
    s = d.system
    assert s.natom == 5 and s.pbtype == systems.PBC_ORBIC and list(s.box) == [20.0, 21.0, 22.0]
    assert np.all(s.eps == 0.238) and np.all(s.sig == 1.705) and np.all(s.x == 0.0)
    assert (d.nsteps, d.save_each, d.seed) == (10, 5, 3)
    bad = [
        _diala_head() + '<movelist><move move_type="rot" move_mode="atom" sel_mode="all"/></movelist>',
        _diala_head().replace('cut_type="Switch"', 'cut_type="shifted"'),
        _diala_head().replace('N="12"', 'N="13"'),
        _diala_head().replace('ens_type="NVT"', 'ens_type="npt"'),
        _diala_head() + '<movelist><move move_type="trn" move_mode="atom" sel_mode="segment"/></movelist>',
    ]
    for k, b in enumerate(bad):
        with pytest.raises(D.DeckError):
            D.read_deck(_write(tmp_path, b, "bad%d.xml" % k))


@pytest.mark.skipif(not os.path.isdir(REF_TEST), reason="the reference tree is not present")
@pytest.mark.parametrize("name", ["argon_box", "diala_vac", "nma_solvated", "trhbn_solvated"])
def test_reference_decks(name, golden, oracle_lib):
    d = D.read_deck(os.path.join(REF_TEST, name + ".xml"))
    key = name if name != "diala_vac" else "diala_vac"
    assert d.system.natom == golden[key]["natom"]
    _check_energy(d.system, golden[key]["energy"])


@pytest.mark.gpu
def test_run_deck_frames(tmp_path):
    """A deck run on the device: the frames land every save_each steps in the
    reference's file formats, and the chain equals the one run without
    frames."""
    from mc_cpp_amd import mcx
    d = D.read_deck(os.path.join(DECKS, "argon_box.xml"))
    d.nsteps, d.save_each, d.seed = 400, 50, 77
    out = D.run_deck(d, out_dir=str(tmp_path))
    lines = (tmp_path / "ener.dat").read_text().splitlines()
    assert [int(l.split("\t")[0]) for l in lines] == list(range(50, 401, 50))
    assert abs(float(lines[-1].split("\t")[1]) - out["etot"]) <= 1e-6 * abs(out["etot"])
    xyz = (tmp_path / "tr.xyz").read_text().splitlines()
    n = d.system.natom
    assert len(xyz) == 8 * (n + 2) and xyz[0] == str(n) and xyz[1] == "#At step 50"
    assert re.fullmatch(r"\S+\t\s*-?\d+\.\d{5}\t\s*-?\d+\.\d{5}\t\s*-?\d+\.\d{5}", xyz[2])
    last = np.array([[float(v) for v in l.split("\t")[1:]] for l in xyz[-n:]])
    # the same chain without frames
    d2 = D.read_deck(os.path.join(DECKS, "argon_box.xml"))
    ff = mcx.FField(d2.system)
    r = ff.h.mc_run(d2.system.moves, 400, 77, rng_mode=mcx.MC_REPLAY, list_every=50)
    assert r["etot"] == out["etot"]
    x, y, z = ff.get_coords()
    assert np.max(np.abs(last - np.stack([x, y, z], 1))) <= 1e-5
