"""The reference's own test systems kept as data fixtures under tests/golden/:
test/diala_vac and test/argon_box as plain text, test/nma_solvated and
test/trhbn_solvated xz-compressed (unpacked into a temporary folder on use).
golden.json records each COR file's SHA-256 and the reference's energies."""
import hashlib
import json
import lzma
import os
import shutil

from mc_cpp_amd import systems

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
GOLD = json.load(open(os.path.join(GOLDEN, "golden.json")))

_FILES = {
    "diala_vac": ("diala_vac/ala2.cor", "diala_vac/ff.dat"),
    "diala_vac_switch": ("diala_vac/ala2.cor", "diala_vac/ff.dat"),
    "argon_box": ("argon_box/ar.cor", "argon_box/forfield.dat"),
    "nma_solvated": ("nma_solvated/nma.cor.xz", "nma_solvated/forfield.dat.xz"),
    "trhbn_solvated": ("trhbn_solvated/trhbn.cor.xz", "trhbn_solvated/forfield.dat.xz"),
}


def _plain(rel, tmpdir):
    src = os.path.join(GOLDEN, rel)
    if not rel.endswith(".xz"):
        return src
    dst = os.path.join(tmpdir, os.path.basename(rel)[:-3])
    with lzma.open(src, "rb") as f, open(dst, "wb") as g:
        shutil.copyfileobj(f, g)
    return dst


def reference_system(name, tmpdir):
    """(System, golden entry) of one of the reference's test systems, read
    from the fixtures with IO_MDBAS's conventions; the COR file is checked
    against the SHA-256 recorded when the goldens were made."""
    g = GOLD[name]
    cor_rel, ff_rel = _FILES[name]
    cor, ff = _plain(cor_rel, str(tmpdir)), _plain(ff_rel, str(tmpdir))
    assert hashlib.sha256(open(cor, "rb").read()).hexdigest() == g["cor_sha256"]
    return systems.read_mdbas(cor, ff, g["box"], pbtype=g["pbtype"], cutmode=g["cutmode"]), g
