"""Force-bias (smart Monte Carlo) translations of the device loop against a
host replay built on the FP64 oracle.

The drift uses beta = 1/(R T) in mol/kcal and F in kcal/mol/A (MV_FBIAS),
or F in iu/A (MV_FBIAS_ACC: the drift of the density the test samples); the
acceptance test applies beta to dE in iu, as the reference's test does for
every move, so force-bias and TRN/ROT moves leave one density invariant (the
Metropolis-Hastings ratio holds for any drift).

Force-bias moves are not in the reference; they follow Rossky, Doll and
Friedman, J. Chem. Phys. 69, 4628 (1978): a moved set's displacement is
dx = beta A F + sqrt(2A) xi (F the net nonbonded force on the set, A the
step in A^2, xi standard normal), accepted with min(1, exp(-beta dE) *
q(new->old) / q(old->new)).  The host replay draws the same
std::mt19937 + generate_canonical numbers as the device (move type,
instance, four uniforms for three Box-Muller normals, the acceptance
uniform only for a negative exponent), takes F from the oracle's FP64
nonbonded gradient (orc_nonbonded_switch_grad) and dE from two oracle
getE calls, so every quantity the device reports is recomputed
independently.

Tolerances (stated per quantity):
  * move instance and accept flag: exact;
  * displacement: 1e-6 A.  The device takes pair displacements from its
    fixed-point coordinate mirror (resolution L/2^32, ~6e-9 A here), which
    moves a pair force by ~1e-8 relative; the drift beta A F then differs
    by ~1e-8 A;
  * ln q ratio: 1e-5 * max(1, |v_fw|^2/4A + |v_rv|^2/4A), the two terms
    of the ratio;
  * dE: 1e-5 * max(|dE|, kT), the per-move bar;
  * final coordinates: 1e-6 A; running total vs a fresh energy: 1e-6
    relative.
"""
import math

import numpy as np
import pytest

from mc_cpp_amd import mcx, systems
from oracle.oracle import MT19937, Oracle, rint

KB_REF = 0.83144621 / 418.4     # beta = 1/(KB_REF*T) per iu (src/MC_metropolis.cpp:175-204)
KCALTOIU = 418.4


def _box(seed=5, nside=8, L=24.85):
    s, _ = systems.water_box(seed=seed, nside=nside, L=L, tag=700 + seed)
    return s


def _fb_move(s, A, typ=systems.MV_FBIAS):
    sele = (np.arange(s.natom) % 3 == 0).astype(np.int32)     # the oxygen of each water
    return systems.MoveType(typ, systems.MODE_RESIDUE, A, 50.0, 0, sele)


def _wrap(v, L):
    return v - L * rint(v / L)


def fb_replay(s, move, nsteps, seed):
    """Host restatement of the force-bias chain on System s (coordinates are
    updated in place); returns the per-step records."""
    o = Oracle(s)
    mt = MT19937(seed)
    beta = 1.0 / (KB_REF * s.temp)
    A = move.dmax
    # drift force in kcal/mol/A (FBIAS) or in iu/A (FBIAS_ACC: the test's own beta)
    fscale = 1.0 if move.type == systems.MV_FBIAS_ACC else 1.0 / KCALTOIU
    atoms = np.nonzero(move.sele)[0]
    out = dict(inst=[], disp=[], de=[], lnq=[], lnq_scale=[], acc=[])

    def net_force(idx):
        o.update_list()
        g = o.grad_switch()
        return -np.array([g[c][idx].sum() for c in range(3)])

    for _ in range(nsteps):
        tp = int(1.0 * mt.canonical())
        assert tp == 0
        inst = int(len(atoms) * mt.canonical())
        idx = np.nonzero(s.resid == s.resid[atoms[inst]])[0]
        u = [mt.canonical() for _ in range(4)]
        a, b = math.sqrt(-2.0 * math.log(1.0 - u[0])), math.sqrt(-2.0 * math.log(1.0 - u[2]))
        xi = np.array([a * math.cos(2 * math.pi * u[1]), a * math.sin(2 * math.pi * u[1]),
                       b * math.cos(2 * math.pi * u[3])])
        o.update_list()
        e0 = o.getE()[0]
        f_old = net_force(idx) * fscale
        disp = beta * A * f_old + math.sqrt(2.0 * A) * xi
        old = [c[idx].copy() for c in (s.x, s.y, s.z)]
        for c, arr in enumerate((s.x, s.y, s.z)):
            arr[idx] = [_wrap(v + disp[c], s.box[c]) for v in old[c]]
        o.update_list()
        e1 = o.getE()[0]
        f_new = net_force(idx) * fscale
        de = e1 - e0
        fw = disp - beta * A * f_old
        rv = -disp - beta * A * f_new
        lnq = (fw @ fw - rv @ rv) / (4.0 * A)
        lnq_scale = (fw @ fw + rv @ rv) / (4.0 * A)
        x = lnq - beta * de
        acc = 1 if x >= 0.0 else int(mt.canonical() < math.exp(x))
        if not acc:
            for c, arr in enumerate((s.x, s.y, s.z)):
                arr[idx] = old[c]
        for k, v in (("inst", inst), ("disp", disp), ("de", de), ("lnq", lnq), ("lnq_scale", lnq_scale),
                     ("acc", acc)):
            out[k].append(v)
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.gpu
@pytest.mark.parametrize("typ,A", [(systems.MV_FBIAS, 1e-3), (systems.MV_FBIAS, 1e-2), (systems.MV_FBIAS_ACC, 5e-6)])
def test_force_bias_replay_matches_host(typ, A):
    """FBIAS_ACC: the drift at the test's own beta (F in iu/A), A ~418x
    smaller for drifts of the same size."""
    s = _box()
    ref = s.copy()
    move = _fb_move(s, A, typ)
    n, seed = 150, 2024
    ff = mcx.FField(s)
    res = ff.h.mc_run([move], n, seed, rng_mode=mcx.MC_REPLAY, list_every=10)
    host = fb_replay(ref, move, n, seed)
    kT = 0.83144621 * s.temp
    assert np.array_equal(res["mvinst"], host["inst"])
    for k in range(n):
        assert np.max(np.abs(res["disp"][k] - host["disp"][k])) <= 1e-6, (k, res["disp"][k], host["disp"][k])
        assert abs(res["lnq"][k] - host["lnq"][k]) <= 1e-5 * max(1.0, host["lnq_scale"][k]), \
            (k, res["lnq"][k], host["lnq"][k])
        assert abs(res["de"][k] - host["de"][k]) <= 1e-5 * max(abs(host["de"][k]), kT), (k, res["de"][k], host["de"][k])
        assert res["acc"][k] == host["acc"][k], k
    # the acceptance uniform was drawn (negative exponent) and both outcomes
    # occur (with the test's own beta the drift follows the force closely
    # enough that this relaxing lattice accepts nearly every move; larger A
    # carries molecules beyond the list skin between the rebuilds)
    x = res["lnq"] - res["de"] / (KB_REF * s.temp)
    if typ == systems.MV_FBIAS:
        assert np.any(x < 0.0) and 0 < int(res["acc"].sum()) < n
    else:
        assert int(res["acc"].sum()) > n // 2
    x, y, z = ff.get_coords()
    assert np.max(np.abs(x - ref.x)) <= 1e-6 and np.max(np.abs(z - ref.z)) <= 1e-6
    e_fresh = ff.h.energy()[0]
    assert abs(res["etot"] - e_fresh) <= 1e-6 * abs(e_fresh)


@pytest.mark.gpu
def test_force_bias_mixed_with_reference_moves():
    """FBIAS next to TRN and ROT residue moves in one Philox chain: records
    are consistent and the running total matches a fresh energy."""
    s = _box(seed=6)
    moves = [s.moves[0], s.moves[1], _fb_move(s, 1e-3)]
    ff = mcx.FField(s)
    res = ff.h.mc_run(moves, 3000, 77, rng_mode=mcx.MC_PHILOX, list_every=50)
    typ = res["mvtyp"]
    assert set(np.unique(typ)) == {0, 1, 2}
    assert np.all(res["disp"][typ == 1] == 0.0) and np.all(res["lnq"][typ != 2] == 0.0)
    assert np.all(np.linalg.norm(res["disp"][typ == 0], axis=1) <= moves[0].dmax * (1 + 1e-12))
    assert np.any(res["lnq"][typ == 2] != 0.0)
    assert int(res["accepted"][2]) > 0
    e_fresh = ff.h.energy()[0]
    assert abs(res["etot"] - e_fresh) <= 1e-6 * abs(e_fresh)


FB_KERNELS = {"one_at_a_time": {"MCX_SPECULATE": "0"},
              "wide_default": {},
              "wide5": {"MCX_WIDE": "5"},
              "wide16": {"MCX_WIDE": "16", "MCX_WIDE_SPLIT": "3"}}


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["fbias", "fbias_acc_adapt", "mixed_philox"])
def test_force_bias_wide_kernel_same_chain(case, monkeypatch):
    """Move sets with force-bias types run speculative rounds on the wide
    kernel: one workgroup per proposal computes the net force on its set,
    the drifted trial, dE and the force at the new position (reduced as the
    one-proposal kernel reduces them), and the rounds resolve in chain order
    with the force-bias acceptance (the acceptance draw taken only for a
    negative exponent).  Every configuration gives the chain of the
    one-proposal kernel (MCX_SPECULATE=0) bit for bit: decisions, dE, ln q,
    displacements, dmax, running total and coordinates (MCX_WIDE_SPLIT is
    ignored: one workgroup per force-bias proposal)."""
    import copy
    s = _box(seed=8)
    if case == "fbias":
        moves, mode = [_fb_move(s, 1e-3)], mcx.MC_REPLAY
    elif case == "fbias_acc_adapt":
        m = _fb_move(s, 5e-6, systems.MV_FBIAS_ACC)
        m.each = 70
        moves, mode = [m], mcx.MC_REPLAY
    else:
        moves, mode = [s.moves[0], s.moves[1], _fb_move(s, 1e-3)], mcx.MC_PHILOX
    n = 2000
    out = {}
    for name, env in FB_KERNELS.items():
        for k in ("MCX_SPECULATE", "MCX_WIDE", "MCX_WIDE_SPLIT"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        sc = copy.deepcopy(s)
        ff = mcx.FField(sc)
        print("case", case, "kernel", name, flush=True)
        res = ff.h.mc_run(copy.deepcopy(moves), n, 31, rng_mode=mode, list_every=50)
        out[name] = (res, ff.get_coords())
    r0, c0 = out["one_at_a_time"]
    assert 0 < int(r0["acc"].sum()) < n
    for name in FB_KERNELS:
        r1, c1 = out[name]
        for key in ("acc", "de", "lnq", "mvtyp", "mvinst", "disp", "trials", "accepted", "dmax"):
            assert np.array_equal(r0[key], r1[key]), (name, key)
        assert r0["etot"] == r1["etot"], name
        for a, b in zip(c0, c1):
            assert np.array_equal(a, b), name
    if case == "fbias_acc_adapt":
        assert r0["dmax"][0] != moves[0].dmax


@pytest.mark.gpu
def test_batch_refuses_force_bias():
    s = _box(seed=7)
    h = mcx.from_system(s)
    with pytest.raises(mcx.MCXError):
        mcx.Batch([h], [_fb_move(s, 1e-4)])


def test_lnq_is_the_gaussian_proposal_log_ratio():
    """The acceptance factor's closed form against scipy's Gaussian log
    densities of the forward and reverse proposals."""
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(3)
    for _ in range(20):
        A, beta = 10.0 ** rng.uniform(-5, -2), rng.uniform(0.5, 3.0)
        f_old, f_new = rng.normal(size=3) * 1e3, rng.normal(size=3) * 1e3
        disp = beta * A * f_old + math.sqrt(2 * A) * rng.normal(size=3)
        cov = 2.0 * A * np.eye(3)
        fwd = multivariate_normal(beta * A * f_old, cov).logpdf(disp)
        rev = multivariate_normal(beta * A * f_new, cov).logpdf(-disp)
        a, b = disp - beta * A * f_old, -disp - beta * A * f_new
        lnq = (a @ a - b @ b) / (4.0 * A)
        assert abs(lnq - (rev - fwd)) <= 1e-9 * max(1.0, abs(rev), abs(fwd))
