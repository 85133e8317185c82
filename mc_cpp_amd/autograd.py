"""torch.autograd bridge to the hand-written adjoint kernel.

`nonbonded_energy(h, types, x, y, z, eps_t, sig_t, q_t, eps14_t, sig14_t)`
returns the nonbonded energy elec + vdw (iu) as a differentiable float64
tensor: in SWITCH mode the list pairs inside the cutoff
(src/FField_MDBAS.cpp:336-424) and the switched 1-4 pairs (:426-488), in
FULL mode every non-excluded pair (:194-265) and the 1-4 pairs (:267-318).
Its backward pass is the device adjoint: dE/dx, dE/dy, dE/dz per atom and
dE/d(eps, sig, q, eps14, sig14) summed per atom type in FP64.  Parameters
are per type (`types[i]` is atom i's type) in the AtomList convention of
src/IO_MDBAS.cpp:460-465: eps = sqrt(eps_kcal * 418.4), sig = sigma / 2;
eps14_t / sig14_t are the 1-4 pairs' own parameters, required when the
system has 1-4 pairs.

No host round trip for coordinates or gradients: the forward pass copies
x, y, z device-to-device into the library's own coordinate arrays (viewed
as torch tensors through mcx_device_coords and the CUDA array interface)
and refreshes the fixed-point mirror on the device (mcx_coords_updated);
the backward pass writes the gradients straight into torch tensors
(mcx_nonbonded_grad_device).  Only the per-type parameters travel through
the host, when they change.  Every evaluation runs on the HIP kernels;
inputs on the CPU are moved to the GPU and results returned on the device
of `x`.
"""
from __future__ import annotations

import numpy as np
import torch



class _DeviceView:
    """n float64 values of device memory, presented to torch.as_tensor
    through the CUDA array interface (no copy)."""

    def __init__(self, ptr: int, n: int):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f8", "data": (int(ptr), False),
                                         "version": 2, "strides": None}


def coordinate_views(h):
    """The handle's FP64 master coordinate arrays as torch tensors on the
    GPU (views of the library's memory, cached on the handle)."""
    v = getattr(h, "_torch_views", None)
    if v is None:
        v = tuple(torch.as_tensor(_DeviceView(h.device_coords(k), h.n), device="cuda") for k in range(3))
        h._torch_views = v
    return v


def _np(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


def _set_parameters(h, ty, eps_t, sig_t, q_t, eps14_t, sig14_t):
    """Per-atom parameters from the per-type tables, uploaded only when
    they differ from the handle's current ones."""
    tabs = tuple(None if t is None else _np(t) for t in (eps_t, sig_t, q_t, eps14_t, sig14_t))

    def same(a, b):
        return (a is None and b is None) or (a is not None and b is not None and np.array_equal(a, b))

    last = getattr(h, "_autograd_params", None)
    if last is not None and np.array_equal(last[0], ty) and all(same(a, b) for a, b in zip(tabs, last[1])):
        return
    e, s, q = tabs[0][ty], tabs[1][ty], tabs[2][ty]
    e14 = tabs[3][ty] if tabs[3] is not None else None
    s14 = tabs[4][ty] if tabs[4] is not None else None
    h.set_atoms(q, e, s, e14, s14)                # mass, frozen and residues are kept
    h._autograd_params = (ty.copy(), tabs)


class NonbondedEnergy(torch.autograd.Function):
    """E_nb(x, y, z, eps_t, sig_t, q_t, eps14_t, sig14_t) on handle `h` (an
    mcx.Handle whose bonded terms and exclusions are set up)."""

    @staticmethod
    def forward(ctx, h, types, x, y, z, eps_t, sig_t, q_t, eps14_t=None, sig14_t=None):
        ty = np.ascontiguousarray(np.asarray(types, dtype=np.int32))
        if len(h.get_exclusions()[2]) and (eps14_t is None or sig14_t is None):
            raise ValueError("this system has 1-4 pairs: pass their per-type eps14_t and sig14_t")
        _set_parameters(h, ty, eps_t, sig_t, q_t, eps14_t, sig14_t)
        vx, vy, vz = coordinate_views(h)
        for v, c in ((vx, x), (vy, y), (vz, z)):
            v.copy_(c.detach().reshape(-1).to(device=v.device, dtype=torch.float64))
        torch.cuda.current_stream(vx.device).synchronize()     # the library runs on its own stream
        h.coords_updated()
        h.update_list()                               # (nothing to build in FULL mode)
        ener = h.energy()
        ctx.h, ctx.ty, ctx.nt = h, ty, int(eps_t.numel())
        ctx.dev = x.device
        ctx.has14 = eps14_t is not None and sig14_t is not None
        return torch.tensor(ener[3] + ener[4], dtype=torch.float64, device=x.device)

    @staticmethod
    def backward(ctx, g):
        h, n, nt = ctx.h, ctx.h.n, ctx.nt
        dev = coordinate_views(h)[0].device
        gpos = torch.empty((3, n), dtype=torch.float64, device=dev)
        gtype = torch.empty((nt, 5), dtype=torch.float64, device=dev)
        h.grad_device(gpos.data_ptr(), ctx.ty, nt, gtype.data_ptr())
        w = g.to(device=dev, dtype=torch.float64)
        gp = (gpos * w).to(ctx.dev)
        gt = (gtype * w).to(ctx.dev)
        g14 = (gt[:, 3], gt[:, 4]) if ctx.has14 else (None, None)
        return (None, None, gp[0], gp[1], gp[2], gt[:, 0], gt[:, 1], gt[:, 2]) + g14


def nonbonded_energy(h, types, x, y, z, eps_t, sig_t, q_t, eps14_t=None, sig14_t=None):
    return NonbondedEnergy.apply(h, types, x, y, z, eps_t, sig_t, q_t, eps14_t, sig14_t)


def atom_types(s):
    """Type index per atom from distinct (charge, eps, sig), and the per-type
    parameter arrays (eps_t, sig_t, q_t) of System s."""
    key = np.stack([s.charge, s.eps, s.sig], 1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    return inv.astype(np.int32).reshape(-1), uniq[:, 1].copy(), uniq[:, 2].copy(), uniq[:, 0].copy()


def atom_types14(s):
    """As atom_types, with the 1-4 parameters in the key: (types, eps_t,
    sig_t, q_t, eps14_t, sig14_t)."""
    key = np.stack([s.charge, s.eps, s.sig, s.eps14, s.sig14], 1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    return (inv.astype(np.int32).reshape(-1), uniq[:, 1].copy(), uniq[:, 2].copy(), uniq[:, 0].copy(),
            uniq[:, 3].copy(), uniq[:, 4].copy())
