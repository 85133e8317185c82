"""The drop-in adapter where the reference would use it: the reference's own
Metropolis loop, MC_metropolis::run (src/MC_metropolis.cpp:45-173), with
FField_MCX (include/mcx_ffield_adapter.hpp) as its FField instead of
FField_MDBAS -- two getE() per move, each uploading the coordinates the
reference's move code changed and evaluating the energy on the GPU, and
askListUpdate() every 50 moves rebuilding the pair list on the GPU.

The loop, moves, random numbers and coordinate backups are the reference's
own code (oracle/_ref/libmcref_mcx.so, built from its sources by
oracle/Makefile target "ref-mcx"); only the force field differs.  The runs
are checked against the reference's recorded outcome (tests/golden) and
against the same loop over FField_MDBAS, getE() call for call.
"""
import hashlib
import json
import os
import tempfile

import numpy as np
import pytest

from mc_cpp_amd import systems
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))

CASES = {
    # 864 Ar atoms, atom translations, 1,000 moves (seed 12345)
    "argon_fcc_seed1": lambda: systems.argon_fcc(seed=1),
    # 512 waters, residue translations and rotations, 300 moves (seed 777)
    "water512_seed3": lambda: systems.water_box(seed=3, nside=8, L=24.85),
}

needs_lib = pytest.mark.skipif(not O.refmcx_available(),
                               reason="oracle/_ref/libmcref_mcx.so not built (needs the reference sources)")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _run(name, gpu):
    s, fp = CASES[name]()
    g = GOLD[name]["mc"]
    d = tempfile.mkdtemp(prefix="mcx_dropin_")
    cor, ff = systems.save_mdbas(os.path.join(d, name), s, fp)
    r = O.ReferenceLoop(cor, ff, s, gpu=gpu)
    res = r.mc_run(s.moves, g["nsteps"], g["seed"])
    x, y, z = r.coords()
    res["xyz"] = np.stack([x, y, z], 1)
    return res, g


def _check_against_golden(res, g):
    assert res["trials"].tolist() == g["trials"]
    assert res["accepted"].tolist() == g["accepted"]
    assert res["dmax"].tolist() == g["dmax"]
    assert sha(res["xyz"]) == g["final_xyz_sha256"]
    # the initial total, then eold and enew of every step
    assert len(res["energies"]) == 1 + 2 * g["nsteps"]


@needs_lib
def test_reference_loop_harness_on_cpu():
    """The recording harness leaves the reference's loop as it is: over
    FField_MDBAS it ends on the reference's own recorded outcome."""
    res, g = _run("argon_fcc_seed1", gpu=False)
    _check_against_golden(res, g)


@needs_lib
@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_reference_loop_over_dropin(name):
    """Same accepted moves, dmax and final coordinates (bit for bit: the
    reference's move code moves the atoms) as the pure-reference run, and
    every getE() within the library's energy tolerance (1e-6 of the total;
    the Lennard-Jones part is summed from FP32 terms) of FField_MDBAS's."""
    res, g = _run(name, gpu=True)
    _check_against_golden(res, g)
    ref, _ = _run(name, gpu=False)
    eg, ec = res["energies"], ref["energies"]
    assert eg.shape == ec.shape
    scale = np.abs(ec).max()
    err = np.abs(eg - ec).max() / scale
    print("max |getE gpu - getE reference| / max |E| = %.3g" % err)
    assert err <= 1e-6
    # the energy changes the acceptance test saw
    dg, dc = eg[2::2] - eg[1::2], ec[2::2] - ec[1::2]
    assert np.abs(dg - dc).max() <= 1e-6 * scale
