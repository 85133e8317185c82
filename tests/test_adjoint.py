"""The adjoint kernel (forces and parameter gradients of the switched
nonbonded energy), its per-type FP64 sums, the batched launch and the
torch.autograd wrapper, on the GPU through the C ABI.

References and tolerances (stated per quantity):
  * FP64 oracle gradient (orc_nonbonded_switch_grad) on 64 seeded water
    boxes: every component within 1e-4 of the largest component of its kind
    (the LJ pair arithmetic is FP32);
  * central finite differences, positions (h = 1e-4 A): each difference is
    taken with the device's per-move dE (FP64 pair changes, ~1e-10 iu), on a
    box whose bonded force constants are zero so that dE is the nonbonded
    change alone: relative 1e-4 of the force scale;
  * central finite differences, parameters (1e-5 relative step): of the
    oracle's FP64 nonbonded energy, per atom type: relative 1e-4;
  * batched launch: bit-identical to one launch per system;
  * autograd: forward equals elec + vdw of getE(ener[10]); backward equals
    the adjoint scaled by the incoming gradient.
"""
import numpy as np
import pytest

from mc_cpp_amd import autograd, mcx, systems
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _box(seed, nside=8, L=24.85):
    s, _ = systems.water_box(seed=seed, nside=nside, L=L, tag=500 + seed)
    return s


def _check_grad(g, o, tag):
    for k, name in enumerate(("gx", "gy", "gz", "geps", "gsig", "gq")):
        scale = max(np.max(np.abs(o[k])), 1e-12)
        err = np.max(np.abs(g[k] - o[k]))
        assert err <= 1e-4 * scale, (tag, name, err, scale)


def test_gradient_vs_fp64_oracle_64_configurations():
    """64 seeded 12,288-atom water boxes (the headline box's size and
    density) in one batched launch, each checked against the FP64 gradient,
    plus the batched launch against per-system launches."""
    boxes = [systems.water_box(seed=100 + k)[0] for k in range(64)]
    handles = [mcx.FField(s).h for s in boxes]
    grads = mcx.grad_batch(handles)
    for k, (s, h, g) in enumerate(zip(boxes, handles, grads)):
        o = Oracle(s.copy())
        o.update_list(16)
        _check_grad(g, o.grad_switch(), "config %d" % k)
        if k % 16 == 0:
            single = h.grad()
            for a, b in zip(g, single):
                assert np.array_equal(a, b)
    del handles


@pytest.mark.parametrize("cutmode", [systems.CUT_SWITCH, systems.CUT_FULL])
def test_gradient_with_14_pairs_and_full_mode(cutmode):
    """The bonded_mix system (orthorhombic box, chains with 1-4 pairs that
    have their own eps14/sig14, waters): the adjoint of elec + vdw in SWITCH
    mode (list pairs + switched 1-4 pairs) and FULL mode (all non-excluded
    pairs + 1-4 pairs) against the FP64 gradient (orc_nonbonded_grad, pinned
    by central differences), including dE/deps14 and dE/dsig14 per atom
    from the device-memory entry point; every component within 1e-4 of the
    largest of its kind.  The water lattice is shaken by up to 0.01 A: on
    the bare lattice some oxygen pairs sit exactly half a box apart, where
    the all-pair energy has a kink (the minimum image switches) and the
    gradient is one-sided."""
    import torch
    s, _ = systems.bonded_mix(seed=7, cutmode=cutmode)
    rng = np.random.default_rng(3)
    for a in (s.x, s.y, s.z):
        a += rng.uniform(-0.01, 0.01, s.natom)
    h = mcx.FField(s).h
    o = Oracle(s.copy())
    if cutmode == systems.CUT_SWITCH:
        o.update_list(16)
    ref = o.grad_all()
    assert np.abs(ref[6]).max() > 0 and np.abs(ref[7]).max() > 0
    _check_grad(h.grad(), ref[:6], "bonded_mix")
    n = s.natom
    gpos = torch.empty((3, n), dtype=torch.float64, device="cuda")
    gtype = torch.empty((n, 5), dtype=torch.float64, device="cuda")
    h.grad_device(gpos.data_ptr(), np.arange(n, dtype=np.int32), n, gtype.data_ptr())
    gp, gt = gpos.cpu().numpy(), gtype.cpu().numpy()
    _check_grad([gp[0], gp[1], gp[2], gt[:, 0], gt[:, 1], gt[:, 2]], ref[:6], "bonded_mix device")
    for c, name in ((3, "geps14"), (4, "gsig14")):
        scale = np.abs(ref[3 + c]).max()
        assert np.abs(gt[:, c] - ref[3 + c]).max() <= 1e-4 * scale, (name, np.abs(gt[:, c] - ref[3 + c]).max(), scale)


def test_gradient_vs_fp64_oracle_water_12288():
    s, _ = systems.water_box(seed=2)
    h = mcx.FField(s).h
    o = Oracle(s.copy())
    o.update_list()
    _check_grad(h.grad(), o.grad_switch(), "water12288")


def test_position_gradient_vs_central_differences():
    s, _ = systems.water_box(seed=2)
    s.bonds["k"][:] = 0.0
    s.angles["k"][:] = 0.0
    ff = mcx.FField(s)
    h = ff.h
    gx, gy, gz, _, _, _ = h.grad()
    g = np.stack([gx, gy, gz], 1)
    rng = np.random.default_rng(11)
    atoms = rng.choice(s.natom, 16, replace=False)
    hstep = 1e-4
    scale = np.max(np.abs(g))
    for i in atoms:
        i = int(i)
        for c in range(3):
            p = np.array([s.x[i], s.y[i], s.z[i]])
            pp, pm = p.copy(), p.copy()
            pp[c] += hstep
            pm[c] -= hstep
            ep = h.delta_energy(i, i, pp[0:1], pp[1:2], pp[2:3])
            em = h.delta_energy(i, i, pm[0:1], pm[1:2], pm[2:3])
            fd = (ep - em) / (2 * hstep)
            assert abs(fd - g[i, c]) <= 1e-4 * scale, (i, c, fd, g[i, c])


def _oracle_nb(s):
    e = Oracle(s).getE()
    return e[3] + e[4]


def test_parameter_gradient_vs_central_differences():
    for seed in (0, 1, 2):
        s = _box(seed)
        types, eps_t, sig_t, q_t = autograd.atom_types(s)
        h = mcx.FField(s).h
        _, _, _, ge, gs, gq = h.grad_types(types, len(eps_t))
        for name, table, gdev in (("eps", eps_t, ge), ("sig", sig_t, gs), ("q", q_t, gq)):
            for t in range(len(table)):
                if table[t] == 0.0:
                    continue          # hydrogens carry no LJ
                step = 1e-5 * abs(table[t])
                vals = []
                for sign in (1, -1):
                    c = s.copy()
                    arr = {"eps": c.eps, "sig": c.sig, "q": c.charge}[name]
                    arr[types == t] = table[t] + sign * step
                    vals.append(_oracle_nb(c))
                fd = (vals[0] - vals[1]) / (2 * step)
                assert abs(fd - gdev[t]) <= 1e-4 * max(abs(fd), 1.0), (seed, name, t, fd, gdev[t])


def test_autograd_function():
    """Forward = elec + vdw of getE(ener[10]) at the tensors' coordinates
    (copied device to device into the library's own arrays); backward = the
    adjoint scaled by the incoming gradient."""
    import torch
    s = _box(7)
    types, eps_t, sig_t, q_t = autograd.atom_types(s)
    h = mcx.FField(s).h
    dev = "cuda"
    x, y, z = (torch.tensor(a, dtype=torch.float64, device=dev, requires_grad=True) for a in (s.x, s.y, s.z))
    e, sg, q = (torch.tensor(a, dtype=torch.float64, device=dev, requires_grad=True) for a in (eps_t, sig_t, q_t))
    E = autograd.nonbonded_energy(h, types, x, y, z, e, sg, q)
    ref = h.energy()
    assert float(E) == pytest.approx(ref[3] + ref[4], rel=1e-12)
    (2.0 * E).backward()
    gx, gy, gz, ge, gs, gq = h.grad_types(types, len(eps_t))
    for t, a in ((x, gx), (y, gy), (z, gz), (e, ge), (sg, gs), (q, gq)):
        assert np.allclose(t.grad.cpu().numpy(), 2.0 * a, rtol=1e-12, atol=0)
    # a small step against the gradient lowers the energy
    with torch.no_grad():
        xs = x - 1e-5 * x.grad
        ys = y - 1e-5 * y.grad
        zs = z - 1e-5 * z.grad
    E2 = autograd.nonbonded_energy(h, types, xs, ys, zs, e.detach(), sg.detach(), q.detach())
    assert float(E2) < float(E)
    # the coordinates the forward pass wrote into the library are the tensors'
    hx, hy, hz = h.get_coords()
    assert np.array_equal(hx, xs.cpu().numpy()) and np.array_equal(hz, zs.cpu().numpy())


def test_autograd_with_14_pairs():
    """The autograd function on a system with 1-4 pairs: their own
    per-type eps14/sig14 are differentiable inputs; forward = elec + vdw;
    the gradients equal the FP64 gradient summed per type (1e-4 of the
    largest of each kind)."""
    import torch
    s, _ = systems.bonded_mix(seed=7)
    types, eps_t, sig_t, q_t, e14_t, s14_t = autograd.atom_types14(s)
    h = mcx.FField(s).h
    o = Oracle(s.copy())
    o.update_list(16)
    ref = o.grad_all()
    xyz = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True) for a in (s.x, s.y, s.z)]
    par = [torch.tensor(a, dtype=torch.float64, device="cuda", requires_grad=True)
           for a in (eps_t, sig_t, q_t, e14_t, s14_t)]
    E = autograd.nonbonded_energy(h, types, *xyz, *par)
    eo = o.getE()
    assert float(E) == pytest.approx(eo[3] + eo[4], rel=1e-6)
    E.backward()
    for c, t in enumerate(xyz):
        scale = np.abs(ref[c]).max()
        assert np.abs(t.grad.cpu().numpy() - ref[c]).max() <= 1e-4 * scale
    nt = len(eps_t)
    for c, t in enumerate(par):
        want = np.bincount(types, weights=ref[3 + c], minlength=nt)
        scale = max(np.abs(want).max(), 1e-12)
        assert np.abs(t.grad.cpu().numpy() - want).max() <= 1e-4 * scale, c
    with pytest.raises(ValueError):
        autograd.nonbonded_energy(h, types, *xyz, *par[:3])
