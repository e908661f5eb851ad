"""Flattening of potential trees into fused-kernel leaves.

Every GPU potential describes itself at time t as a list of *leaves*: one
native component (an analytic type with its parameters, or an expansion with
the device address of its table descriptor) plus the affine map and weight
under which the wrappers around it evaluate it:

    Phi(x) = sum_leaves A * Phi_leaf(sigma * (x - tau))

Composite concatenates its children's leaves, Shifted adds its center to tau,
Scaled multiplies sigma by 1/scale and A by ampl/scale, Evolving multiplies A
by the lerp weight of each snapshot.  The fused kernel (agx_composite_eval)
then evaluates all leaves of a potential in one launch.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib


@dataclass
class Leaf:
    type: int
    params: Tuple[float, ...] = ()
    table: int = 0                       # device address of a descriptor image
    keepalive: Optional[object] = None   # object owning the device memory behind `table`
    ampl: float = 1.0
    sigma: float = 1.0
    tau: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    prog_flags: int = 0                  # Multipole: agx_multipole.prog_flags of the table

    def shifted(self, c) -> "Leaf":
        return replace(self, tau=(self.tau[0] + c[0], self.tau[1] + c[1], self.tau[2] + c[2]))

    def scaled(self, s: float, a: float) -> "Leaf":
        return replace(self, sigma=self.sigma * s, ampl=self.ampl * a * s,
                       tau=(self.tau[0] / s, self.tau[1] / s, self.tau[2] / s))

    def weighted(self, w: float) -> "Leaf":
        return replace(self, ampl=self.ampl * w)


def to_components(leaves: Sequence[Leaf]):
    arr = (_lib.agx_component * len(leaves))()
    for i, lf in enumerate(leaves):
        c = arr[i]
        c.type = lf.type
        c.prog_flags = lf.prog_flags
        for k, v in enumerate(lf.params):
            c.p[k] = float(v)
        c.table = lf.table or None
        c.ampl = lf.ampl
        c.scale = lf.sigma
        for k in range(3):
            c.center[k] = lf.tau[k]
    return arr


def eval_leaves(leaves: Sequence[Leaf], xyz: torch.Tensor, phi, force, deriv, device) -> None:
    """One fused launch per AGX_MAX_COMPONENTS leaves (accumulating when more)."""
    L = _lib.load()
    n = xyz.shape[0]
    stream = _lib.stream_handle(device)
    what = (_lib.AGX_PHI if phi is not None else 0) | (_lib.AGX_FORCE if force is not None else 0) | \
           (_lib.AGX_DERIV if deriv is not None else 0)
    chunks = [leaves[i:i + _lib.AGX_MAX_COMPONENTS]
              for i in range(0, len(leaves), _lib.AGX_MAX_COMPONENTS)]
    for ci, chunk in enumerate(chunks):
        comps = to_components(chunk)
        if ci == 0:
            outs = (phi, force, deriv)
        else:
            outs = tuple(None if o is None else torch.empty_like(o) for o in (phi, force, deriv))
        rc = L.agx_composite_eval(comps, len(chunk), xyz.data_ptr(), n,
                                  None if outs[0] is None else outs[0].data_ptr(),
                                  None if outs[1] is None else outs[1].data_ptr(),
                                  None if outs[2] is None else outs[2].data_ptr(), what, stream)
        _lib.check(rc, "agx_composite_eval")
        if ci > 0:
            for acc, part in zip((phi, force, deriv), outs):
                if acc is not None:
                    acc.add_(part)


def density_leaves(leaves: Sequence[Leaf], xyz: torch.Tensor, rho: torch.Tensor, device) -> None:
    L = _lib.load()
    n = xyz.shape[0]
    stream = _lib.stream_handle(device)
    for i in range(0, len(leaves), _lib.AGX_MAX_COMPONENTS):
        chunk = leaves[i:i + _lib.AGX_MAX_COMPONENTS]
        out = rho if i == 0 else torch.empty_like(rho)
        rc = L.agx_composite_density(to_components(chunk), len(chunk), xyz.data_ptr(), n,
                                     out.data_ptr(), stream)
        _lib.check(rc, "agx_composite_density")
        if i > 0:
            rho.add_(out)
