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
    """64 seeded 1,536-atom boxes in one batched launch, checked against the
    FP64 gradient, plus the batched launch against per-system launches."""
    boxes = [_box(seed) for seed in range(64)]
    handles = [mcx.FField(s).h for s in boxes]
    grads = mcx.grad_batch(handles)
    for k, (s, h, g) in enumerate(zip(boxes, handles, grads)):
        o = Oracle(s.copy())
        o.update_list()
        _check_grad(g, o.grad_switch(), "config %d" % k)
        if k % 16 == 0:
            single = h.grad()
            for a, b in zip(g, single):
                assert np.array_equal(a, b)


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
    import torch
    s = _box(7)
    types, eps_t, sig_t, q_t = autograd.atom_types(s)
    h = mcx.FField(s).h
    dev = "cuda" if torch.cuda.is_available() else "cpu"
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
