#!/usr/bin/env python3
"""Writes tests/golden/golden.json: reference outputs for the parity tests.

Every number here comes from the mc_cpp reference's own code, compiled from
its sources into oracle/_ref/libmcref.so (oracle/Makefile, target "ref") and
driven through oracle/ref_driver.cpp, on
  * the reference's own test inputs (test/diala_vac, test/argon_box, copied
    here as data; test/nma_solvated and test/trhbn_solvated read in place);
  * seeded synthetic systems from mc_cpp_amd.systems written as MDBAS files.
Run in the build container (it needs /root/reference):  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mc_cpp_amd import systems  # noqa: E402
from oracle import oracle  # noqa: E402

REF = "/root/reference/test"


def h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def list_digest(r):
    cnt, rows = r.list_rows()
    flat = np.concatenate([x for x in rows if len(x)] or [np.zeros(0, np.int32)]).astype(np.int32)
    return {"npairs": int(cnt.sum()), "counts_sha256": h(cnt.astype(np.int32)), "flat_sha256": h(flat)}


def main():
    oracle.build(ref=True)
    out = {"units": "energies in iu (1 kcal/mol = 418.4 iu), order tot,pot,kin,elec,vdw,bond,ang,ub,dihe,impr"}
    tmp = tempfile.mkdtemp()

    # reference test systems
    cases = [("diala_vac", "diala_vac/ala2.cor", "diala_vac/ff.dat", [30, 30, 30], 12, 0),
             ("diala_vac_switch", "diala_vac/ala2.cor", "diala_vac/ff.dat", [30, 30, 30], 12, 1),
             ("argon_box", "argon_box/ar.cor", "argon_box/forfield.dat", [30, 30, 30], 256, 1),
             ("nma_solvated", None, None, None, 2658, 0), ("trhbn_solvated", None, None, None, 21977, 0)]
    for name, cor, ff, box, n, cut in cases:
        if cor is None:
            xml = open(os.path.join(REF, name + ".xml")).read()
            import re
            a = re.search(r'pbc_type="(\w+)"\s+a="([\d.]+)"\s+b="([\d.]+)"\s+c="([\d.]+)"', xml)
            ptype = {"cubic": 1, "orbic": 2}[a.group(1).lower()]
            box = [float(a.group(2)), float(a.group(3)), float(a.group(4))]
            cor = os.path.join(REF, re.search(r'cor_path="([^"]+)"', xml).group(1))
            ff = os.path.join(REF, re.search(r'ff_file="([^"]+)"', xml).group(1))
        else:
            ptype = 1
            cor, ff = os.path.join(HERE, cor), os.path.join(HERE, ff)
        s = systems.read_mdbas(cor, ff, box, pbtype=ptype, cutmode=cut)
        r = oracle.Reference(cor, ff, s)
        e = r.getE()
        rows, l14 = r.excl_rows()
        ent = {"box": box, "pbtype": ptype, "cutmode": cut, "natom": n, "energy": e.tolist(),
               "cor_sha256": hashlib.sha256(open(cor, "rb").read()).hexdigest(),
               "excl_counts_sha256": h(np.array([len(x) for x in rows], np.int32)),
               "excl_flat_sha256": h(np.concatenate([x for x in rows if len(x)] or [np.zeros(0, np.int32)]).astype(np.int32)),
               "pairs14": l14.tolist() if len(l14) < 64 else h(l14.astype(np.int32))}
        if cut == 1:
            ent["list"] = list_digest(r)
        out[name] = ent
        print(name, e[0] / 418.4)

    # synthetic systems
    synth = {"argon_fcc_seed1": systems.argon_fcc(seed=1),
             "water512_seed3": systems.water_box(seed=3, nside=8, L=24.85),
             "water4096_seed2": systems.water_box(seed=2),
             # a box smaller than twice the cutoff (replica-ladder sizes)
             "water90_small_box": systems.water_box(seed=21, nside=5, L=(90 / 0.0334) ** (1 / 3), nmol=90,
                                                    temp=290.0, tag=300),
             # orthorhombic box with harmonic and Morse bonds, Urey-Bradley,
             # cosine and harmonic dihedrals and impropers, 1-4 pairs
             "bonded_mix_seed7": systems.bonded_mix(seed=7)}
    for name, (s, fp) in synth.items():
        cor, ff = systems.save_mdbas(os.path.join(tmp, name), s, fp)
        r = oracle.Reference(cor, ff, s)
        ent = {"energy": r.getE().tolist(), "list": list_digest(r)}
        s2 = s.copy(); s2.cutmode = systems.CUT_FULL
        cor2, ff2 = systems.save_mdbas(os.path.join(tmp, name + "_full"), s2, fp)
        ent["energy_full"] = oracle.Reference(cor2, ff2, s2).getE().tolist()
        if name in ("argon_fcc_seed1", "water512_seed3", "water90_small_box"):
            nsteps = 1000 if name.startswith("argon") else 300
            seed = 12345 if name.startswith("argon") else 11 if name == "water90_small_box" else 777
            mc = r.mc_run(s.moves, nsteps, seed)
            x, y, z = r.coords()
            ent["mc"] = {"nsteps": nsteps, "seed": seed, "trials": mc["trials"].tolist(),
                         "accepted": mc["accepted"].tolist(), "dmax": mc["dmax"].tolist(),
                         "final_xyz_sha256": h(np.stack([x, y, z], 1))}
        out[name] = ent
        print(name, ent["energy"][0] / 418.4)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
