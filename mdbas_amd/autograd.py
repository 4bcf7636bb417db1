"""torch.autograd front end of the C-ABI adjoint (mdg_backward,
mdg_backward_traj).

Leaves are FP64 torch tensors: per-atom charges and per-type LJ
parameters in the units of a FORF file (eps in kcal/mol, Rmin/2 in A).
They map onto the per-atom arrays the kernels use exactly as the reference
reads them: eps_i = sqrt(eps_kcal * 418.4) (io.c:1174) and
sig_i = (Rmin/2) / 2^(1/6) (tools/prepare.c:1999-2000), so torch carries
the chain rule from the per-atom gradients the library returns to the
leaves.  The energy and its gradient both come from the HIP library; torch
only composes them.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Context, MDG_FP64

SIXTH_ROOT_2 = 2.0 ** (1.0 / 6.0)


def per_atom_lj(eps_kcal: torch.Tensor, rmin_half: torch.Tensor, types: torch.Tensor):
    """Per-type (eps kcal/mol, Rmin/2 A) -> the per-atom eps, sig arrays."""
    eps = torch.sqrt(eps_kcal[types] * 418.4)
    sig = rmin_half[types] / SIXTH_ROOT_2
    return eps, sig


class Water:
    """One device context plus the initial state the functions below start
    from (FP64)."""

    def __init__(self, wb, precision: int = MDG_FP64):
        self.wb = wb
        self.ctx = Context(wb, precision=precision)
        self.x0 = [a.copy() for a in (wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)]
        self.types = torch.as_tensor((wb.eps == 0).astype(np.int64))   # 0 = O, 1 = H

    def _set(self, q, eps, sig):
        self.ctx.set_params(q.detach().cpu().numpy(), eps.detach().cpu().numpy(),
                            sig.detach().cpu().numpy())
        self.ctx.upload(*self.x0)

    def close(self):
        self.ctx.close()


class _Potential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, eps, sig, water):
        water._set(q, eps, sig)
        e = water.ctx.initial_energy().pot
        ctx.water = water
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        d = ctx.water.ctx.backward(float(g))
        t = [torch.from_numpy(d[k].copy()) for k in range(3)]
        return t[0], t[1], t[2], None


class _TimeAveraged(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, eps, sig, water, nsteps, ckpt):
        water._set(q, eps, sig)
        water.ctx.initial_energy()
        J, d = water.ctx.backward_traj(nsteps, ckpt)
        ctx.grad = d
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        d = ctx.grad * float(g)
        t = [torch.from_numpy(d[k].copy()) for k in range(3)]
        return t[0], t[1], t[2], None, None, None


def potential_energy(water: Water, q, eps_kcal, rmin_half):
    """E_pot of the initial configuration (internal units, 10 J/mol)."""
    eps, sig = per_atom_lj(eps_kcal, rmin_half, water.types)
    return _Potential.apply(q, eps, sig, water)


def time_averaged_potential(water: Water, q, eps_kcal, rmin_half, nsteps: int, ckpt: int):
    """Mean E_pot over steps 1..nsteps of NVE + SHAKE from the initial
    state; the gradient comes from the checkpointed reverse sweep."""
    eps, sig = per_atom_lj(eps_kcal, rmin_half, water.types)
    return _TimeAveraged.apply(q, eps, sig, water, nsteps, ckpt)
