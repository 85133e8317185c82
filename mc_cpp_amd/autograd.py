"""torch.autograd bridge to the hand-written adjoint kernel.

`nonbonded_energy(h, types, x, y, z, eps_t, sig_t, q_t)` returns the
switched nonbonded energy elec + vdw of src/FField_MDBAS.cpp:336-424 (iu)
as a differentiable float64 tensor.  Its backward pass is the device
adjoint (mcx_nonbonded_grad_types): forces dE/dx, dE/dy, dE/dz per atom and
dE/d(eps, sig, q) summed per atom type in FP64.  Parameters are per type
(`types[i]` is atom i's type) in the AtomList convention of
src/IO_MDBAS.cpp:460-465: eps = sqrt(eps_kcal * 418.4), sig = sigma / 2.

The energy and its gradient are evaluated by the HIP kernels only; tensors
travel through host memory to the C ABI, so the function accepts CPU or
GPU tensors and returns results on the device of `x`.  Systems with 1-4
pairs are refused: their 1-4 terms use separate parameters the adjoint
does not cover.
"""
from __future__ import annotations

import numpy as np
import torch

from . import mcx


def _np(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class NonbondedEnergy(torch.autograd.Function):
    """E_nb(x, y, z, eps_t, sig_t, q_t) on handle `h` (an mcx.Handle whose
    bonded terms and exclusions are set up)."""

    @staticmethod
    def forward(ctx, h, types, x, y, z, eps_t, sig_t, q_t):
        ty = np.ascontiguousarray(np.asarray(types, dtype=np.int32))
        if len(h.get_exclusions()[2]):
            raise ValueError("the adjoint covers list pairs only; this system has 1-4 pairs")
        e, s, q = _np(eps_t)[ty], _np(sig_t)[ty], _np(q_t)[ty]
        h.set_atoms(q, e, s)                      # mass, frozen and residues are kept
        h.set_coords(_np(x), _np(y), _np(z))
        h.update_list()
        ener = h.energy()
        ctx.h, ctx.ty, ctx.nt = h, ty, int(eps_t.numel())
        ctx.dev = x.device
        return torch.tensor(ener[3] + ener[4], dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, g):
        gx, gy, gz, ge, gs, gq = ctx.h.grad_types(ctx.ty, ctx.nt)
        w = float(g)
        out = [torch.from_numpy(a * w).to(ctx.dev) for a in (gx, gy, gz, ge, gs, gq)]
        return (None, None) + tuple(out)


def nonbonded_energy(h, types, x, y, z, eps_t, sig_t, q_t):
    return NonbondedEnergy.apply(h, types, x, y, z, eps_t, sig_t, q_t)


def atom_types(s):
    """Type index per atom from distinct (charge, eps, sig), and the per-type
    parameter arrays (eps_t, sig_t, q_t) of System s."""
    key = np.stack([s.charge, s.eps, s.sig], 1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    return inv.astype(np.int32).reshape(-1), uniq[:, 1].copy(), uniq[:, 2].copy(), uniq[:, 0].copy()
