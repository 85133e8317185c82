#!/usr/bin/env python3
"""Writes tests/golden/replay_water12288.npz: a 10,000-move Metropolis
replay of the 12,288-atom water box (4,096 waters, L = 49.7 A, switch
10/12 A, skin 2 A, T = 300 K, residue TRN 0.3 A and ROT 15 deg, pair list
rebuilt every 50 moves, std::mt19937 seed 20261204) computed by the FP64
oracle (oracle/libmc_oracle.c, MC_metropolis::run with the moving-set dE,
which takes the decisions of the reference's two full getE per move).

Stored: per step the move type, instance, dE (iu) and accept flag; the
final dmax per move type, total energy change and the final coordinates of
every 64th atom.  The oracle's energy sums run serially (the reference's
order); its list builds use all cores (the rows do not depend on it).
Run anywhere the oracle builds:  python tests/golden/make_replay_water12288.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mc_cpp_amd import systems  # noqa: E402
from oracle import oracle  # noqa: E402

NSTEPS = 10000
SEED = 20261204
EVERY = 50


def main():
    oracle.build(ref=False)
    s, _ = systems.water_box(seed=2)
    o = oracle.Oracle(s)
    threads = max(1, len(os.sched_getaffinity(0)))
    o.update_list(threads)
    e0 = o.getE()[0]
    res = o.mc_run(NSTEPS, SEED, mode=1, list_every=EVERY, nthreads=threads)
    pick = np.arange(0, s.natom, 64)
    np.savez_compressed(os.path.join(HERE, "replay_water12288.npz"),
                        nsteps=np.int64(NSTEPS), seed=np.int64(SEED), list_every=np.int64(EVERY),
                        mvtyp=res["mvtyp"].astype(np.int8), mvinst=res["mvinst"].astype(np.int32),
                        de=res["de"], acc=res["acc"].astype(np.uint8), dmax=res["dmax"],
                        de_total=np.float64(res["etot"] - e0), pick=pick.astype(np.int32),
                        xyz=np.stack([s.x[pick], s.y[pick], s.z[pick]], 1))
    print("accepted", int(res["acc"].sum()), "of", NSTEPS, "dE total", res["etot"] - e0)


if __name__ == "__main__":
    main()
