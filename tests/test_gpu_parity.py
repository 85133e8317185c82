"""Parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances (stated per quantity):
  * total energy: relative 1e-6 (FP32 pair arithmetic, FP64 sums);
  * elec / vdw components: 1e-6 of max(|component|, |total|);
  * bonded components (FP64 on the device): relative 1e-9;
  * pair list: identical sets, rows in the same ascending order (bit-exact);
  * per-move dE: |dE_gpu - dE_fp64| <= 1e-5 * max(|dE|, kT) with
    kT = 0.83144621 * T iu (0.596 kcal/mol at 300 K);
  * replayed Metropolis runs: identical accept decisions except a move
    whose dE lies within the dE tolerance of the acceptance threshold
    (after such a move the two chains legitimately diverge).
"""
import json
import os

import numpy as np
import pytest

from mc_cpp_amd import mcx, systems
from oracle.oracle import Oracle

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")))

pytestmark = pytest.mark.gpu


def _pair(s):
    ref = s.copy()
    o = Oracle(ref)
    return o, ref


def _check_energy(eg, eo, tag):
    tot = abs(eo[0])
    assert abs(eg[0] - eo[0]) <= 1e-6 * tot, (tag, "tot", eg[0], eo[0])
    for k, name in ((3, "elec"), (4, "vdw")):
        assert abs(eg[k] - eo[k]) <= 1e-6 * max(abs(eo[k]), tot), (tag, name, eg[k], eo[k])
    for k, name in ((5, "bond"), (6, "ang"), (7, "ub"), (8, "dihe"), (9, "impr")):
        assert abs(eg[k] - eo[k]) <= 1e-9 * max(abs(eo[k]), 1e-6), (tag, name, eg[k], eo[k])


def _check_list(h, o):
    cnt_o, off_o, flat_o = o.list_csr()
    cnt_g, flat_g = h.get_list()
    assert np.array_equal(cnt_o, cnt_g)
    assert np.array_equal(flat_o, flat_g)
    return int(cnt_g.sum())


SYSTEMS = {
    "argon864": lambda: systems.argon_fcc(seed=1)[0],
    "water1536": lambda: systems.water_box(seed=3, nside=8, L=24.85)[0],
    "mixed_small": lambda: systems.mixed_system(seed=3, nside=8, L=24.85, nbead=60)[0],
    # orthorhombic; harmonic and Morse bonds, Urey-Bradley, cosine and
    # harmonic dihedrals and impropers, 1-4 pairs with their own parameters
    "bonded_mix": lambda: systems.bonded_mix(seed=7)[0],
}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_switch_energy_and_list(name):
    s = SYSTEMS[name]()
    o, ref = _pair(s)
    o.update_list()
    eo = o.getE()
    ff = mcx.FField(s)
    npairs = _check_list(ff.h, o)
    assert npairs > 0
    _check_energy(ff.getE10(), eo, name)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_full_energy(name):
    s = SYSTEMS[name]()
    s.cutmode = systems.CUT_FULL
    o, ref = _pair(s)
    eo = o.getE()
    ff = mcx.FField(s)
    _check_energy(ff.getE10(), eo, name + "/full")


def test_bonded_mix_against_reference_goldens():
    """Every bonded kind against the reference's own numbers (the mc_cpp
    sources built in oracle/_ref, recorded in golden.json), switched and
    all-pair, in the orthorhombic box."""
    g = GOLD["bonded_mix_seed7"]
    for cut, key in ((systems.CUT_SWITCH, "energy"), (systems.CUT_FULL, "energy_full")):
        s = systems.bonded_mix(seed=7)[0]
        s.cutmode = cut
        e = mcx.FField(s).getE10()
        _check_energy(e, np.array(g[key]), "bonded_mix/" + key)
        assert abs(e[5]) > 1e4 and abs(e[7]) > 1e3 and abs(e[8]) > 1e4 and abs(e[9]) > 1e4


@pytest.mark.parametrize("name", ["nma_solvated", "trhbn_solvated"])
def test_reference_solvated_systems(name, tmp_path):
    """The reference's test/nma_solvated (2,658 atoms, cubic) and
    test/trhbn_solvated (21,977 atoms, orthorhombic, 1,777 Urey-Bradley
    terms), all-pair as their decks run them, against the reference's
    energies; then switched (10/12 A, 2 A skin) against the FP64 oracle."""
    from tests.refdata import reference_system
    s, g = reference_system(name, tmp_path)
    assert s.cutmode == systems.CUT_FULL
    e = mcx.FField(s.copy()).getE10()
    _check_energy(e, np.array(g["energy"]), name + "/full")
    if name == "trhbn_solvated":
        assert len(s.ub["a"]) == 1777 and abs(e[7]) > 0.0
    s.cutmode = systems.CUT_SWITCH
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    _check_list(ff.h, o)
    _check_energy(ff.getE10(), o.getE(), name + "/switch")


def test_morse_bond_closed_form():
    """A single Morse bond on the device (src/FField_MDBAS.cpp:519-527):
    E = D (1 - e^{-b (d - r0)})^2, zero at r0."""
    D, b, r0 = 80.0 * 418.4, 1.8, 1.53
    for d in (1.53, 1.2, 1.9, 3.0):
        s = systems.System(natom=2, box=[50.0] * 3, pbtype=systems.PBC_NONE, cutmode=systems.CUT_FULL,
                           x=[0.0, d], y=[0.0, 0.0], z=[0.0, 0.0], resid=[1, 1])
        s.bonds = systems.make_terms("bonds", a=[0], b=[1], type=[1], k=[D], r0=[r0], beta=[b])
        e = mcx.FField(s).getE10()
        want = D * (1.0 - np.exp(-b * (d - r0))) ** 2
        assert abs(e[5] - want) <= 1e-12 * D, (d, e[5], want)
        assert abs(e[0] - want) <= 1e-12 * D


def test_water_box_12288_energy_and_list():
    """Config-2 box at full size: 12,288 atoms, 7.07 M list pairs."""
    s, _ = systems.water_box(seed=2)
    o, ref = _pair(s)
    o.update_list()
    eo = o.getE()
    ff = mcx.FField(s)
    n = _check_list(ff.h, o)
    assert 6.5e6 < n < 7.5e6
    _check_energy(ff.getE10(), eo, "water12288")
    s2 = s.copy()
    s2.cutmode = systems.CUT_FULL
    o2 = Oracle(s2)
    ff2 = mcx.FField(s2)
    _check_energy(ff2.getE10(), o2.getE(), "water12288/full")


def test_list_after_coordinate_update_and_boundaries():
    """Atoms placed exactly at +-L/2 and pairs at the list radius."""
    s, _ = systems.argon_fcc(seed=9, ncell=3, L=17.2)
    L = s.box[0]
    s.x[0], s.y[0], s.z[0] = L / 2, -L / 2, 0.0
    s.x[1], s.y[1], s.z[1] = -L / 2 + 14.0, -L / 2, 0.0   # exactly at ctoff + dcut from atom 0 through the image
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    _check_list(ff.h, o)
    # move every atom a little and rebuild through askListUpdate
    rng = np.random.default_rng(3)
    for arr in (s.x, s.y, s.z):
        arr += rng.uniform(-0.4, 0.4, s.natom)
    ref.x[:], ref.y[:], ref.z[:] = s.x, s.y, s.z
    ff.set_coords(s.x, s.y, s.z)
    ff.askListUpdate(50)
    o.update_list()
    _check_list(ff.h, o)


@pytest.mark.parametrize("box", ["cubic", "orthorhombic"])
def test_list_pairs_at_the_list_radius(box):
    """Pairs placed at (ctoff + dcut) * (1 + e) for e from -4e-6 to +4e-6
    (inside and outside the band the FP32 candidate test hands to the FP64
    test) and at a few ulps from it; the list must equal the oracle's."""
    s, _ = systems.argon_fcc(seed=12)
    if box == "orthorhombic":
        s.pbtype = systems.PBC_ORBIC
        s.box = np.array([34.39, 36.1, 33.2])
    rc = s.ctoff + s.dcut
    rng = np.random.default_rng(17)
    rel = [-4e-6, -3.9e-6, -1e-6, -1e-7, -1e-12, -1e-15, 0.0, 1e-15, 1e-12, 1e-7, 1e-6, 3.9e-6, 4e-6, 5e-6]
    for k in range(200):
        a, b = 4 * k, 4 * k + 1
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        d = rc * (1.0 + rel[k % len(rel)]) * u
        pos = np.array([s.x[a], s.y[a], s.z[a]]) + d
        pos -= s.box * np.trunc(pos / s.box + np.where(pos >= 0, 0.5, -0.5))
        s.x[b], s.y[b], s.z[b] = pos
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    _check_list(ff.h, o)


LIST_CASES = {
    "argon864": lambda: systems.argon_fcc(seed=1)[0],
    "water1536": lambda: systems.water_box(seed=3, nside=8, L=24.85)[0],
    "water_small_box": lambda: systems.water_box(seed=21, nside=5, L=(90 / 0.0334) ** (1 / 3), nmol=90)[0],
    "bonded_mix": lambda: systems.bonded_mix(seed=7)[0],
    "bonded_mix_vacuum": lambda: systems.bonded_mix(seed=7, pbtype=systems.PBC_NONE)[0],
    "argon_small_box": lambda: systems.argon_fcc(seed=9, ncell=3, L=17.2)[0],
}


@pytest.mark.parametrize("group", ["16", "5", "32"])
@pytest.mark.parametrize("name", list(LIST_CASES))
def test_sorted_candidate_list_build(name, group, monkeypatch):
    """The sorted-candidate list kernel (taken by default from 32,768 atoms
    up, here forced with MCX_LIST_SORTED=1 at every size and group size):
    the group's candidates in ascending atom order from one bitmap, hit masks
    per candidate, ordered row output.  The list must equal the oracle's bit
    for bit: cubic, orthorhombic and vacuum boxes, boxes too small for the
    cell-centre prefilter, frozen atoms, exclusions and 1-4 partners."""
    monkeypatch.setenv("MCX_LIST_SORTED", "1")
    monkeypatch.setenv("MCX_LIST_GROUP", group)
    s = LIST_CASES[name]()
    if name == "argon864":
        s.frozen[: s.natom // 3] = 1
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    print("case", name, "group", group, flush=True)
    _check_list(ff.h, o)
    # after coordinates move: a rebuild through askListUpdate
    rng = np.random.default_rng(5)
    for arr in (s.x, s.y, s.z):
        arr += rng.uniform(-0.3, 0.3, s.natom)
    ref.x[:], ref.y[:], ref.z[:] = s.x, s.y, s.z
    ff.set_coords(s.x, s.y, s.z)
    ff.askListUpdate(50)
    o.update_list()
    _check_list(ff.h, o)


def test_frozen_pairs_excluded():
    s, _ = systems.argon_fcc(seed=4)
    s.frozen[: s.natom // 2] = 1
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    _check_list(ff.h, o)


def _kT(s):
    return 0.83144621 * s.temp


@pytest.mark.parametrize("name", ["argon864", "water1536", "mixed_small", "bonded_mix", "bonded_mix_beads",
                                  "bonded_mix_full"])
def test_delta_energy(name):
    """Residue translations (bonded_mix_beads: single chain beads, whose
    bonds, angles, Urey-Bradley terms, torsions and 1-4 pairs change;
    bonded_mix_full: all-pair in the orthorhombic box)."""
    s = SYSTEMS[name.replace("_beads", "").replace("_full", "")]()
    if name.endswith("_full"):
        s.cutmode = systems.CUT_FULL
    o, ref = _pair(s)
    if s.cutmode == systems.CUT_SWITCH:
        o.update_list()
    ff = mcx.FField(s)
    rng = np.random.default_rng(11)
    x, y, z = ref.x, ref.y, ref.z
    beads = np.nonzero(np.asarray(s.resnames) == "CHN")[0] if name.startswith("bonded") else None
    for trial in range(30):
        if name.endswith("_beads"):
            a = b = int(rng.choice(beads))
        else:
            i = int(rng.choice(beads)) if (beads is not None and trial % 2) else int(rng.integers(0, s.natom))
            r = s.resid[i]
            idx = np.nonzero(s.resid == r)[0]
            a, b = int(idx[0]), int(idx[-1])
        d = rng.uniform(-0.3, 0.3, 3)
        nx, ny, nz = x[a:b + 1] + d[0], y[a:b + 1] + d[1], z[a:b + 1] + d[2]
        de_g = ff.h.delta_energy(a, b, nx, ny, nz)
        e_old = o.moving_set_energy(a, b)
        sx, sy, sz = x[a:b + 1].copy(), y[a:b + 1].copy(), z[a:b + 1].copy()
        x[a:b + 1], y[a:b + 1], z[a:b + 1] = nx, ny, nz
        e_new = o.moving_set_energy(a, b)
        x[a:b + 1], y[a:b + 1], z[a:b + 1] = sx, sy, sz
        de_o = e_new - e_old
        tol = 1e-5 * max(abs(de_o), _kT(s))
        assert abs(de_g - de_o) <= tol, (name, trial, a, b, de_g, de_o)


def _compare_replay(res, ro, beta, tol_fn):
    acc_g, acc_o = res["acc"], ro["acc"]
    head = [(int(res["mvtyp"][k]), int(ro["mvtyp"][k]), int(res["mvinst"][k]), int(ro["mvinst"][k]),
             float(res["de"][k]), float(ro["de"][k])) for k in range(min(5, len(acc_g)))]
    print("first steps (typ gpu/fp64, inst gpu/fp64, dE gpu/fp64):", head)
    mism = np.nonzero(acc_g != acc_o)[0]
    n_ok = len(acc_g) if len(mism) == 0 else int(mism[0])
    bad = np.nonzero((res["mvinst"] != ro["mvinst"]) | (res["mvtyp"] != ro["mvtyp"]))[0]
    if len(bad):
        print("first differing move pick at step", int(bad[0]))
    dd = np.nonzero(np.abs(res["de"] - ro["de"]) > np.array([tol_fn(v) for v in ro["de"]]))[0]
    if len(dd):
        k = int(dd[0])
        print("first dE outside tolerance at step", k, "pick gpu", int(res["mvtyp"][k]), int(res["mvinst"][k]),
              "fp64", int(ro["mvtyp"][k]), int(ro["mvinst"][k]), "acc before", int(acc_g[:k].sum()), int(acc_o[:k].sum()))
    # dE agreement on the common prefix
    for k in range(n_ok):
        assert abs(res["de"][k] - ro["de"][k]) <= tol_fn(ro["de"][k]), (k, res["de"][k], ro["de"][k])
    if len(mism):
        k = int(mism[0])
        # a legitimate divergence: the two dE straddle the acceptance threshold
        assert abs(res["de"][k] - ro["de"][k]) <= tol_fn(ro["de"][k]), (k, res["de"][k], ro["de"][k])
    return n_ok


def test_mc_replay_argon_1e5():
    """1e5 replayed single-atom moves on the config-1 argon box."""
    s, _ = systems.argon_fcc(seed=1)
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    n = 100000
    res = mcx.MC_metropolis(ff, nsteps=n, seed=20240601).run()
    ro = o.mc_run(n, 20240601, mode=1)
    kT = _kT(s)
    n_ok = _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT))
    assert n_ok == n, "accept sequence diverged at step %d" % n_ok
    assert np.array_equal(res["trials"], np.bincount(ro["mvtyp"], minlength=1))
    x, y, z = ff.get_coords()
    assert np.max(np.abs(x - ref.x)) < 1e-9


@pytest.mark.parametrize("which", ["trn", "rot", "both"])
def test_mc_replay_water_residue_moves(which):
    s, _ = systems.water_box(seed=3, nside=8, L=24.85)
    if which == "trn":
        s.moves = s.moves[:1]
    elif which == "rot":
        s.moves = s.moves[1:]
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    n = 3000
    res = mcx.MC_metropolis(ff, nsteps=n, seed=777).run()
    ro = o.mc_run(n, 777, mode=1)
    kT = _kT(s)
    n_ok = _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT))
    assert n_ok == n
    # running total matches a fresh evaluation of the final state
    assert abs(res["etot"] - ff.getE()) <= 1e-6 * abs(ff.getE())


def test_mc_replay_small_box():
    """Half the box (6.95 A) is shorter than the cutoff: every pair is in the
    list and many moves change a pair's minimum image."""
    s, _ = systems.water_box(seed=21, nside=5, L=(90 / 0.0334) ** (1 / 3), nmol=90, temp=290.0, tag=300)
    o, ref = _pair(s)
    o.update_list()
    e0 = abs(o.getE()[0])        # the totals' 1e-6 bar applies to the starting energy too
    ff = mcx.FField(s)
    n = 2000
    res = mcx.MC_metropolis(ff, nsteps=n, seed=11).run()
    ro = o.mc_run(n, 11, mode=1)
    kT = _kT(s)
    n_ok = _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT))
    assert n_ok == n
    # running totals agree (each starts from its own full energy); the
    # device's fresh total agrees with the FP64 one
    assert abs(res["etot"] - ro["etot"]) <= 1e-6 * max(e0, abs(ro["etot"]))
    assert abs(ff.getE() - Oracle(ref).getE()[0]) <= 1e-6 * max(e0, abs(ro["etot"]))


KERNELS = {"one_at_a_time": {"MCX_SPECULATE": "0"},
           "read_ahead": {"MCX_SPECULATE": "1", "MCX_WIDE": "0"},
           "wide_default": {"MCX_SPECULATE": "1"},
           "wide5_split1": {"MCX_SPECULATE": "1", "MCX_WIDE": "5", "MCX_WIDE_SPLIT": "1"},
           "wide4_split3": {"MCX_SPECULATE": "1", "MCX_WIDE": "4", "MCX_WIDE_SPLIT": "3"},
           "wide16_split2": {"MCX_SPECULATE": "1", "MCX_WIDE": "16", "MCX_WIDE_SPLIT": "2"}}


def _chain_case(case):
    if case.startswith("water_full"):
        s, _ = systems.water_box(seed=5, nside=5, L=15.53, nmol=125, tag=41)
        s.cutmode = systems.CUT_FULL
    elif case.startswith("water_small_box"):
        s, _ = systems.water_box(seed=21, nside=5, L=(90 / 0.0334) ** (1 / 3), nmol=90, temp=290.0, tag=300)
    elif case.startswith("water"):
        s, _ = systems.water_box(seed=3, nside=8, L=24.85)
    elif case.startswith("bonded"):
        s, _ = systems.bonded_mix(seed=7)
    else:
        s, _ = systems.argon_fcc(seed=1)
    for k, m in enumerate(s.moves):
        m.each = 70 + 40 * k if case.endswith("adapt") else 0
    return s


@pytest.mark.parametrize("case", ["water_every50_adapt", "water_auto_philox", "argon_philox_adapt",
                                  "bonded_every50_adapt", "water_full_cubic", "water_small_box_adapt"])
def test_mc_chain_kernels_same_chain(case, monkeypatch):
    """Move sets without force-bias moves run a speculative kernel: the
    read-ahead kernel (one workgroup, the next proposal drawn and loaded
    while the current move's pairs are evaluated; MCX_WIDE=0) or, by
    default, the wide kernel (MCX_WIDE proposals per round, each some moves
    ahead on the assumption that the moves before it are rejected, its
    pairs shared out over MCX_WIDE_SPLIT workgroups, resolved in chain
    order).  MCX_SPECULATE=0 selects the
    one-proposal-at-a-time kernel.  All give the same chain bit for bit: the
    same dE, decisions, dmax and coordinates, through list rebuilds, chunk
    boundaries (the draw state carried over), dmax adaptation steps (which
    end a round), the displacement-triggered list (an accepted move past
    half the skin ends the chunk), the all-pair instance and a box smaller
    than twice the cutoff."""
    import copy
    s = _chain_case(case)
    mode = mcx.MC_REPLAY if "philox" not in case else mcx.MC_PHILOX
    every = mcx.LIST_AUTO if "auto" in case else 50
    n = 3000
    out = {}
    for name, env in KERNELS.items():
        for k in ("MCX_SPECULATE", "MCX_WIDE", "MCX_WIDE_SPLIT"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        sc = copy.deepcopy(s)
        ff = mcx.FField(sc)
        print("case", case, "kernel", name, flush=True)
        res = ff.h.mc_run(sc.moves, n, 97, rng_mode=mode, list_every=every)
        out[name] = (res, ff.get_coords())
    r0, c0 = out["one_at_a_time"]
    assert 0 < int(r0["acc"].sum()) < n
    for name in KERNELS:
        r1, c1 = out[name]
        for key in ("acc", "de", "mvtyp", "mvinst", "disp", "trials", "accepted", "dmax"):
            assert np.array_equal(r0[key], r1[key]), (name, key)
        assert r0["etot"] == r1["etot"], name
        for a, b in zip(c0, c1):
            assert np.array_equal(a, b), name
    if case.endswith("adapt"):
        assert not np.array_equal(r0["dmax"], np.array([m.dmax for m in s.moves]))


@pytest.mark.parametrize("cut", ["switch", "full"])
def test_mc_replay_bonded_mix(cut):
    """2,000 replayed moves on the orthorhombic system with every bonded
    kind: water residue TRN/ROT, single chain beads (Morse and harmonic
    bonds, angles, Urey-Bradley, both torsion kinds, 1-4 pairs change) and
    whole-chain rotations (16 moved atoms)."""
    s = systems.bonded_mix(seed=7)[0]
    if cut == "full":
        s.cutmode = systems.CUT_FULL
    o, ref = _pair(s)
    if s.cutmode == systems.CUT_SWITCH:
        o.update_list()
    e0 = abs(o.getE()[0])
    ff = mcx.FField(s)
    n = 2000 if cut == "switch" else 800
    res = mcx.MC_metropolis(ff, nsteps=n, seed=4711).run()
    ro = o.mc_run(n, 4711, mode=1)
    kT = _kT(s)
    n_ok = _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT))
    assert n_ok == n
    acc_by_type = np.bincount(ro["mvtyp"], weights=ro["acc"], minlength=4)
    assert acc_by_type[2] > 0, "no chain-bead move accepted: the bonded terms' commit path was not exercised"
    assert abs(res["etot"] - ro["etot"]) <= 1e-6 * max(e0, abs(ro["etot"]))
    x, y, z = ff.get_coords()
    assert max(np.max(np.abs(x - ref.x)), np.max(np.abs(y - ref.y)), np.max(np.abs(z - ref.z))) < 1e-9


def _replay_system(kind):
    if kind == "cubic_full":
        s, _ = systems.water_box(seed=5, nside=5, L=15.53, nmol=125, tag=41)
        s.cutmode = systems.CUT_FULL
    elif kind == "orthorhombic_switch":
        s, _ = systems.argon_fcc(seed=13)
        s.pbtype = systems.PBC_ORBIC
        s.box = np.array([34.39, 36.1, 33.2])
        s.x, s.y, s.z = systems.wrap(s, s.x, s.y, s.z)
    else:   # vacuum_full: a flexible dipeptide, no periodic box, atom moves
        from mc_cpp_amd import deck
        s = deck.read_deck(os.path.join(os.path.dirname(__file__), "golden", "decks", "diala_vac.xml")).system
        s.pbtype = systems.PBC_NONE
    return s


@pytest.mark.parametrize("kind", ["cubic_full", "orthorhombic_switch", "vacuum_full"])
def test_mc_replay_box_and_cutoff_kinds(kind):
    """The chain kernel is compiled per box type and cutoff mode (cubic
    switched, cubic all-pair, and a generic instance): replay each against
    the FP64 oracle."""
    s = _replay_system(kind)
    o, ref = _pair(s)
    if s.cutmode == systems.CUT_SWITCH:
        o.update_list()
    e0 = abs(o.getE()[0])
    ff = mcx.FField(s)
    n = 1500
    res = mcx.MC_metropolis(ff, nsteps=n, seed=31).run()
    ro = o.mc_run(n, 31, mode=1)
    kT = _kT(s)
    n_ok = _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT))
    assert n_ok == n
    assert abs(res["etot"] - ro["etot"]) <= 1e-6 * max(e0, abs(ro["etot"]))


def test_mc_replay_through_list_overflow():
    """A list capacity equal to the largest row of the first build (the
    reference's sizeList is max_i neighPair * 1.04 + 1): rows grow during
    the run, a scheduled rebuild inside mcx_mc_run overflows, the chunks
    queued after it leave the chain untouched, and the library grows the
    list and resumes there.  The chain still equals the oracle's."""
    s, _ = systems.water_box(seed=3, nside=8, L=24.85)
    o, ref = _pair(s)
    o.update_list()
    cnt, rows = o.list_rows()
    full = cnt.astype(np.int64).copy()        # full row = half row + the rows naming i
    for r in rows:
        np.add.at(full, r, 1)
    m0 = int(full.max())
    ff = mcx.FField(s)
    assert ff.h.list_capacity(m0) == m0
    n = 2000
    res = mcx.MC_metropolis(ff, nsteps=n, seed=777).run()
    ro = o.mc_run(n, 777, mode=1)
    kT = _kT(s)
    n_ok = _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT))
    assert n_ok == n
    assert ff.h.list_capacity() > m0, "no rebuild outgrew the capacity: the overflow path was not exercised"


def _full_row_max(o):
    cnt, rows = o.list_rows()
    full = cnt.astype(np.int64).copy()        # full row = half row + the rows naming i
    for r in rows:
        np.add.at(full, r, 1)
    return int(full.max())


def test_mc_final_rebuild_overflow():
    """The run ends on a scheduled rebuild that overflows the capacity:
    with seed 777 on this box the oracle's rows first outgrow the first
    build's largest row at the rebuild of step 300, and no chunk follows it.
    The library must grow the list there, so that the energy asked for
    afterwards sums complete rows and equals the oracle's."""
    s, _ = systems.water_box(seed=3, nside=8, L=24.85)
    o, ref = _pair(s)
    o.update_list()
    m0 = _full_row_max(o)
    ff = mcx.FField(s)
    assert ff.h.list_capacity(m0) == m0
    n = 300
    res = mcx.MC_metropolis(ff, nsteps=n, seed=777).run()
    ro = o.mc_run(n, 777, mode=1)
    kT = _kT(s)
    assert _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT)) == n
    o.update_list()
    assert _full_row_max(o) > m0, "the last rebuild does not outgrow the capacity"
    assert ff.h.list_capacity() > m0
    _check_list(ff.h, o)
    _check_energy(ff.getE10(), o.getE(), "after the final rebuild")


def test_mc_frames_through_list_overflow():
    """Frames every 100 steps with rebuilds every 50 and a rebuild that
    overflows at step 300 (the chunks after it return untouched and the run
    resumes there): every frame is written once, in step order, and the
    chain equals the frameless one."""
    s, _ = systems.water_box(seed=3, nside=8, L=24.85)
    o, ref = _pair(s)
    o.update_list()
    m0 = _full_row_max(o)
    e0 = o.getE()[0]
    ff = mcx.FField(s)
    ff.h.list_capacity(m0)
    steps, etots = [], []
    n = 1000
    res = ff.h.mc_run(s.moves, n, 777, list_every=50, save_each=100,
                      on_frame=lambda st, e, x, y, z: (steps.append(st), etots.append(e)))
    assert steps == list(range(100, n + 1, 100)), steps
    assert ff.h.list_capacity() > m0, "the overflow path was not exercised"
    ro = o.mc_run(n, 777, mode=1)
    kT = _kT(s)
    assert _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT)) == n
    run = e0 + np.cumsum(np.where(ro["acc"] == 1, ro["de"], 0.0))
    for st, e in zip(steps, etots):
        assert abs(e - run[st - 1]) <= 1e-6 * abs(e0), (st, e, run[st - 1])


def test_gradient_vs_fp64_and_fd():
    s, _ = systems.water_box(seed=3, nside=8, L=24.85)
    o, ref = _pair(s)
    o.update_list()
    ff = mcx.FField(s)
    gx, gy, gz, ge, gs, gq = ff.h.grad()
    ox, oy, oz, oe, os_, oq = o.grad_switch()
    scale = np.max(np.abs(ox))
    for g_, o_ in ((gx, ox), (gy, oy), (gz, oz)):
        assert np.max(np.abs(g_ - o_)) <= 1e-4 * scale
    for g_, o_ in ((ge, oe), (gs, os_), (gq, oq)):
        assert np.max(np.abs(g_ - o_)) <= 1e-4 * max(np.max(np.abs(o_)), 1e-12)


@pytest.mark.parametrize("case", ["water1536", "bonded_mix", "water_small_box", "mixed_small_frozen", "water12288",
                                  "water12288_every2500", "water12288_long_rows"])
def test_incremental_list_equals_full_rebuild(case, monkeypatch):
    """Scheduled rebuilds inside mcx_mc_run are incremental updates (rows of
    the atoms moved since the last update rebuilt, the other rows patched
    with sorted edits); MCX_LIST_PATCH=0 rebuilds in full.  Both give the
    same chain bit for bit, and the list after the run equals the FP64
    oracle's list at the final coordinates (bit-exact: the same pairs in the
    same order).  water_small_box has a single cell; mixed_small_frozen has
    frozen atoms; water12288 runs long enough for more than 1,536 atoms to
    move since the binning, which asks for a full build in between;
    water12288_every2500 updates every 2,500 moves, so that one update takes
    more than 1,024 moved atoms (more than one per thread of the row kernel,
    more than one per block of its grid) and as many stale ones;
    water12288_long_rows has a 16 A cutoff (18 A list radius), so that rows
    hold more than 2,048 atoms (more than the row kernels read ahead or
    merge in one pass)."""
    import copy
    if case == "water1536":
        s = systems.water_box(seed=3, nside=8, L=24.85)[0]
    elif case == "bonded_mix":
        s = systems.bonded_mix(seed=7)[0]
    elif case == "water_small_box":
        s = systems.water_box(seed=21, nside=5, L=(90 / 0.0334) ** (1 / 3), nmol=90, temp=290.0, tag=300)[0]
    elif case == "mixed_small_frozen":
        s = systems.mixed_system(seed=3, nside=8, L=24.85, nbead=60)[0]
        s.frozen[::7] = 1
    elif case == "water12288_long_rows":
        s = systems.water_box(seed=2)[0]
        s.ctoff, s.cton = 16.0, 14.0
    else:
        s = systems.water_box(seed=2)[0]
    n = {"water12288": 12000, "water12288_every2500": 5000}.get(case, 3000)
    every = 2500 if case == "water12288_every2500" else 50
    out = {}
    for patch in ("0", "1"):
        monkeypatch.setenv("MCX_LIST_PATCH", patch)
        sc = copy.deepcopy(s)
        ff = mcx.FField(sc)
        print("case", case, "patch", patch, flush=True)
        res = ff.h.mc_run(sc.moves, n, 4242, list_every=every)
        out[patch] = (res, ff.get_coords(), ff.h.get_list())
    (r0, c0, l0), (r1, c1, l1) = out["0"], out["1"]
    assert 0 < int(r0["acc"].sum()) < n
    for key in ("acc", "de", "mvtyp", "mvinst", "trials", "accepted", "dmax"):
        assert np.array_equal(r0[key], r1[key]), key
    assert r0["etot"] == r1["etot"]
    for a, b in zip(c0, c1):
        assert np.array_equal(a, b)
    assert np.array_equal(l0[0], l1[0]) and np.array_equal(l0[1], l1[1])
    # n is a multiple of the update interval: the list was brought up to date at the end
    ref = s.copy()
    ref.x[:], ref.y[:], ref.z[:] = c1
    o = Oracle(ref)
    o.update_list()
    cnt_o, off_o, flat_o = o.list_csr()
    assert np.array_equal(cnt_o, l1[0]) and np.array_equal(flat_o, l1[1])


def test_mc_replay_water12288_10k_against_fixture():
    """The headline workload at its own size: 10,000 replayed Metropolis
    moves on the 12,288-atom water box with the list updated every 50
    moves, against the FP64 oracle's run of the same draws
    (tests/golden/replay_water12288.npz, made by make_replay_water12288.py):
    the same move picks and accept decisions at every step, dE within
    1e-5 of max(|dE|, kT) (the bar of every replay here), the same final
    dmax, total energy change and sampled final coordinates."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "replay_water12288.npz"))
    s, _ = systems.water_box(seed=2)
    ff = mcx.FField(s)
    n = int(g["nsteps"])
    res = ff.h.mc_run(s.moves, n, int(g["seed"]), rng_mode=mcx.MC_REPLAY, list_every=int(g["list_every"]))
    ro = {"acc": g["acc"], "de": g["de"], "mvtyp": g["mvtyp"].astype(np.int32), "mvinst": g["mvinst"]}
    kT = _kT(s)
    n_ok = _compare_replay(res, ro, None, lambda de: 1e-5 * max(abs(de), kT))
    assert n_ok == n, "accept sequence diverged at step %d" % n_ok
    assert np.array_equal(res["mvtyp"], ro["mvtyp"]) and np.array_equal(res["mvinst"], ro["mvinst"])
    assert np.allclose(res["dmax"], g["dmax"], rtol=1e-12, atol=0)
    e0 = float(g["de_total"])
    assert abs(float(res["de"][res["acc"] == 1].sum()) - e0) <= 1e-6 * abs(e0)
    x, y, z = ff.get_coords()
    p = g["pick"]
    assert np.max(np.abs(np.stack([x[p], y[p], z[p]], 1) - g["xyz"])) < 1e-9


def test_energy_over_the_list_of_the_last_update():
    """mcx_set_coords_moved (what the drop-in getE uploads after the
    reference's move code changed AtomList) keeps the pair list of the last
    update, as the reference's getE does between askListUpdate calls: on
    the 12,288-atom box with ten pairs of distant molecules swapped (each
    carried far beyond the 2 A skin, no overlaps) the energy is the FP64
    energy over the old list; mcx_set_coords rebuilds, and the two differ."""
    s, _ = systems.water_box(seed=2)
    o, ref = _pair(s)
    o.update_list(16)
    ff = mcx.FField(s)                   # its list: built at these coordinates
    rng = np.random.default_rng(5)
    for m1, m2 in rng.permutation(s.natom // 3)[:20].reshape(10, 2):
        d = np.array([ref.x[3 * m2] - ref.x[3 * m1], ref.y[3 * m2] - ref.y[3 * m1], ref.z[3 * m2] - ref.z[3 * m1]])
        for a in range(3):
            i, j = 3 * m1 + a, 3 * m2 + a
            ref.x[i] += d[0]; ref.y[i] += d[1]; ref.z[i] += d[2]
            ref.x[j] -= d[0]; ref.y[j] -= d[1]; ref.z[j] -= d[2]
    ff.h.set_coords_moved(ref.x, ref.y, ref.z)
    eg = ff.h.energy()
    eo = o.getE()                        # the oracle's list is the old one too
    _check_energy(eg, eo, "list of the last update")
    ff.h.set_coords(ref.x, ref.y, ref.z)
    eg2 = ff.h.energy()
    o.update_list(16)
    eo2 = o.getE()
    _check_energy(eg2, eo2, "rebuilt list")
    assert eo[0] != eo2[0]


def test_moved_coordinate_uploads_back_to_back():
    """mcx_set_coords_moved stages x, y, z in one pinned buffer and makes the
    device copy the master: two uploads with no energy call between them
    leave the second (the staging buffer is not overwritten while a copy
    from it is in flight); the coordinates read back are the uploaded ones
    bit for bit; re-setting the atom parameters afterwards (which rebuilds
    the fixed-point mirror from the coordinates) keeps them; every energy
    equals the FP64 energy of those coordinates over the list of the last
    update."""
    s, _ = systems.water_box(seed=3, nside=8, L=24.85)
    o, ref = _pair(s)
    o.update_list(16)
    ff = mcx.FField(s)
    rng = np.random.default_rng(11)
    moved = []
    for k in range(2):
        c = [a + rng.uniform(-0.3, 0.3, a.shape) for a in (s.x, s.y, s.z)]
        moved.append(c)
        ff.h.set_coords_moved(*c)
    for _ in range(2):
        ref.x[:], ref.y[:], ref.z[:] = moved[1]
        _check_energy(ff.h.energy(), o.getE(), "second of two uploads")
    x, y, z = ff.h.get_coords()
    for a, b in zip((x, y, z), moved[1]):
        assert np.array_equal(a, b)
    ff.h.set_coords_moved(*moved[0])
    ff.h.set_atoms(s.charge, s.eps, s.sig, s.eps14, s.sig14, s.mass, s.frozen, s.resid)
    ref.x[:], ref.y[:], ref.z[:] = moved[0]
    _check_energy(ff.h.energy(), o.getE(), "after the parameters were set again")
    x, y, z = ff.h.get_coords()
    for a, b in zip((x, y, z), moved[0]):
        assert np.array_equal(a, b)
