"""Displacement-triggered Verlet lists (list_every = MCX_LIST_AUTO) and the
~100k-atom mixed system, through the C ABI.

With MCX_LIST_AUTO the device loop rebuilds the list after any accepted move
that carries an atom more than dcut/2 from its position at the last build,
and evaluates any trial that does so over every atom.  Every dE then sums
exactly the pairs inside the cutoff: the oracle's list-free mode (oracle
mode 2, pinned against brute force in tests/test_oracle.py) is the
reference, with the same std::mt19937 draws.  The reference itself
rebuilds on a fixed 50-move schedule (src/MC_metropolis.cpp:151-154).

Tolerances: per-move dE within 1e-5 * max(|dE|, kT) (the per-move bar);
accept sequences identical except a move whose two dE straddle the
threshold; running total vs a fresh energy within 1e-6 relative; lists
bit-exact.
"""
import numpy as np
import pytest

from mc_cpp_amd import mcx, systems
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _kT(s):
    return 0.83144621 * s.temp


def _tol(s):
    kT = _kT(s)
    return lambda de: 1e-5 * max(abs(de), kT)


def _compare(res, ro, tol):
    acc_g, acc_o = res["acc"], ro["acc"]
    mism = np.nonzero(acc_g != acc_o)[0]
    n_ok = len(acc_g) if len(mism) == 0 else int(mism[0])
    for k in range(n_ok):
        assert abs(res["de"][k] - ro["de"][k]) <= tol(ro["de"][k]), (k, res["de"][k], ro["de"][k])
    if len(mism):
        k = int(mism[0])
        assert abs(res["de"][k] - ro["de"][k]) <= tol(ro["de"][k]), ("divergence outside the band", k)
    return n_ok


def _check_list(h, o):
    cnt_o, off_o, flat_o = o.list_csr()
    cnt_g, flat_g = h.get_list()
    assert np.array_equal(cnt_o, cnt_g)
    assert np.array_equal(flat_o, flat_g)


@pytest.mark.parametrize("name", ["water1536", "mixed_small", "water_small_box", "bonded_mix_large_steps"])
def test_auto_list_matches_exact_cutoff_oracle(name):
    if name == "bonded_mix_large_steps":
        # orthorhombic, every bonded kind, residue and atom moves large
        # enough that many trials leave the list's validity (the filtered
        # scan of every atom)
        s, _ = systems.bonded_mix(seed=7)
        for m, d in zip(s.moves, (1.2, 40.0, 0.9, 20.0)):
            m.dmax = d
        n = 2000
    elif name == "water1536":
        # steps large enough that accepted moves outrun half the skin
        s, _ = systems.water_box(seed=3, nside=8, L=24.85)
        s.moves = [systems.MoveType(systems.MV_TRN, systems.MODE_RESIDUE, 1.2, 50.0, 0),
                   systems.MoveType(systems.MV_ROT, systems.MODE_RESIDUE, 40.0, 50.0, 0)]
        n = 3000
    elif name == "mixed_small":
        s, _ = systems.mixed_system(seed=3, nside=8, L=24.85, nbead=60)
        n = 2000
    else:   # half the box is shorter than the cutoff
        s, _ = systems.water_box(seed=21, nside=5, L=(90 / 0.0334) ** (1 / 3), nmol=90, temp=290.0, tag=300)
        n = 2000
    ref = s.copy()
    ff = mcx.FField(s)
    res = ff.h.mc_run(s.moves, n, 4242, rng_mode=mcx.MC_REPLAY, list_every=mcx.LIST_AUTO)
    o = Oracle(ref)
    o.update_list()
    ro = o.mc_run(n, 4242, mode=2)
    assert _compare(res, ro, _tol(s)) == n
    if name == "water1536":
        assert res["list_updates"] > 0
    x, y, z = ff.get_coords()
    assert np.max(np.abs(x - ref.x)) < 1e-9 and np.max(np.abs(z - ref.z)) < 1e-9
    e_fresh = ff.h.energy()[0]
    assert abs(res["etot"] - e_fresh) <= 1e-6 * abs(e_fresh)
    # the list in force after the run is the reference's list of its build
    # positions, and covers every pair of the final positions
    o.update_list()
    ff.h.update_list()
    _check_list(ff.h, o)


def test_mixed_100k_energy_and_list():
    """The ~100k-atom mixed system: 32,768-site water lattice minus the waters
    within 2.6 A of a 2,000-bead chain (92,384 atoms), L = 99.4 A."""
    s, _ = systems.mixed_system()
    assert s.natom > 90000
    ref = s.copy()
    ff = mcx.FField(s)
    o = Oracle(ref)
    o.update_list()
    _check_list(ff.h, o)
    eg, eo = ff.getE10(), o.getE()
    tot = abs(eo[0])
    assert abs(eg[0] - eo[0]) <= 1e-6 * tot, (eg[0], eo[0])
    for k in (3, 4):
        assert abs(eg[k] - eo[k]) <= 1e-6 * max(abs(eo[k]), tot), (k, eg[k], eo[k])
    for k in (5, 6, 7, 8, 9):
        assert abs(eg[k] - eo[k]) <= 1e-9 * max(abs(eo[k]), 1e-6), (k, eg[k], eo[k])


def test_mixed_100k_auto_list_run():
    """20,000 mixed moves (residue TRN/ROT of waters, atom TRN of the chain)
    with displacement-triggered rebuilds: the running total matches a fresh
    energy and the final list is the oracle's."""
    s, _ = systems.mixed_system()
    ff = mcx.FField(s)
    res = ff.h.mc_run(s.moves, 20000, 99, rng_mode=mcx.MC_PHILOX, list_every=mcx.LIST_AUTO, record=False)
    e_fresh = ff.h.energy()[0]
    assert abs(res["etot"] - e_fresh) <= 1e-6 * abs(e_fresh)
    x, y, z = ff.get_coords()
    ref = s.copy()
    ref.x[:], ref.y[:], ref.z[:] = x, y, z
    o = Oracle(ref)
    o.update_list()
    ff.h.update_list()
    _check_list(ff.h, o)
