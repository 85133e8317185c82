"""Host-side system model: generators, MDBAS writer/reader round trips."""
import os

import numpy as np

from mc_cpp_amd import systems


def test_mdbas_round_trip(tmp_path):
    s, fp = systems.mixed_system(seed=3, nside=6, L=18.6, nbead=30)
    cor, ff = systems.save_mdbas(str(tmp_path / "mixed"), s, fp)
    r = systems.read_mdbas(cor, ff, s.box, temp=s.temp)
    for f in ("x", "y", "z", "charge", "eps", "sig", "eps14", "sig14", "mass", "frozen", "resid"):
        assert np.array_equal(getattr(r, f), getattr(s, f)), f
    for kind in ("bonds", "ub", "angles", "dihedrals", "impropers"):
        a, b = getattr(r, kind), getattr(s, kind)
        for k in a:
            assert np.array_equal(a[k], b[k]), (kind, k)


def test_generators_are_deterministic():
    a, _ = systems.water_box(seed=11, nside=5, L=15.5)
    b, _ = systems.water_box(seed=11, nside=5, L=15.5)
    c, _ = systems.water_box(seed=12, nside=5, L=15.5)
    assert np.array_equal(a.x, b.x) and not np.array_equal(a.x, c.x)


def test_water_geometry_and_charges():
    s, _ = systems.water_box(seed=2)
    assert s.natom == 12288 and abs(s.charge.sum()) < 1e-9
    x = s.coords.reshape(-1, 3, 3)
    L = s.box[0]
    for h in (1, 2):
        d = x[:, h] - x[:, 0]
        d -= L * np.round(d / L)
        assert np.allclose(np.linalg.norm(d, axis=1), 0.9572, atol=1e-12)
    assert np.all(np.abs(s.coords) <= L / 2 + 1e-12)


def test_argon_density_and_jitter():
    s, _ = systems.argon_fcc(seed=1)
    assert s.natom == 864
    assert abs(s.natom / s.box.prod() - 0.02125) < 1e-4


def test_mixed_chain_is_neutral_and_clash_free():
    s, _ = systems.mixed_system(seed=3, nside=8, L=24.85, nbead=60)
    chain = slice(s.natom - 60, s.natom)
    assert abs(s.charge[chain].sum()) < 1e-12
    assert len(s.dihedrals["a"]) == 57 and len(s.bonds["a"]) == (s.natom - 60) + 59


def test_replica_ladder():
    reps = systems.water_replicas(seed=4, nrep=8)
    T = [r[0].temp for r in reps]
    assert abs(T[0] - 280.0) < 1e-9 and abs(T[-1] - 400.0) < 1e-9
    for s, _ in reps:
        nmol = s.natom // 3
        assert 300 <= nmol <= 3000 and abs(nmol / s.box.prod() - 0.0334) < 1e-9
