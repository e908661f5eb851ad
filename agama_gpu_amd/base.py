"""Common Agama-style API shared by every GPU potential.

Mirrors the reference's per-class methods (gpu_potential.py:826-935,
_analytic_potentials.py:95-160) once, on top of a single primitive each
class implements:

    _evaluate(xyz, want_phi, want_force, want_deriv, t) -> (phi, force, deriv)
    _density(xyz, t) -> rho

where ``xyz`` is a contiguous (N,3) float64 tensor on the potential's device
and unrequested outputs are None.  Inputs may be torch tensors (any device)
or numpy arrays; outputs are torch tensors on the GPU.  A (3,) input returns
a scalar / (3,) / (6,) result, as the reference's squeeze rule
(gpu_potential.py:167-176).
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from ._lib import require_gpu

Tensor = torch.Tensor


def prep_xyz(xyz, device: torch.device) -> Tuple[Tensor, bool]:
    """(3,) or (N,3) array -> contiguous float64 (N,3) tensor on ``device``."""
    if isinstance(xyz, torch.Tensor):
        t = xyz
    else:
        t = torch.as_tensor(np.asarray(xyz, dtype=np.float64))
    if t.dtype != torch.float64:
        t = t.to(torch.float64)
    if t.dim() == 1:
        if t.shape[0] != 3:
            raise ValueError(f"Single-point xyz must have shape (3,), got {tuple(t.shape)}")
        return t.reshape(1, 3).to(device).contiguous(), True
    if t.dim() != 2 or t.shape[1] != 3:
        raise ValueError(f"xyz must have shape (N,3), got {tuple(t.shape)}")
    return t.to(device).contiguous(), False


def _sq(a: Optional[Tensor], single: bool):
    if a is None:
        return None
    return a[0] if single else a


class PotentialBase:
    """Base of all GPU potentials: API methods and ``+`` composition."""

    def __init__(self, device=None):
        self.device = norm_device(device) if device is not None else require_gpu()

    # ---- primitives -----------------------------------------------------------
    def _leaves(self, t=0.0):
        """This potential at time t as fused-kernel leaves (see leaves.py)."""
        raise NotImplementedError

    def _evaluate(self, xyz: Tensor, want_phi: bool, want_force: bool, want_deriv: bool,
                  t=0.0, out=None):
        """Phi / force / deriv at xyz; ``out``: optional preallocated
        (phi, force, deriv) tensors (None entries are allocated).  Default: one
        fused launch over this potential's leaves."""
        from .leaves import eval_leaves
        phi, force, deriv = alloc_outputs(xyz.shape[0], self.device, want_phi, want_force,
                                          want_deriv, out)
        eval_leaves(self._leaves(t), xyz, phi, force, deriv, self.device)
        return phi, force, deriv

    def _density(self, xyz: Tensor, t=0.0) -> Tensor:
        from .leaves import density_leaves
        rho = torch.empty(xyz.shape[0], dtype=torch.float64, device=self.device)
        density_leaves(self._leaves(t), xyz, rho, self.device)
        return rho

    # ---- Agama API ----------------------------------------------------------
    def potential(self, xyz, t=0.0):
        x, single = prep_xyz(xyz, self.device)
        phi, _, _ = self._evaluate(x, True, False, False, t)
        return _sq(phi, single)

    def force(self, xyz, t=0.0):
        x, single = prep_xyz(xyz, self.device)
        _, f, _ = self._evaluate(x, False, True, False, t)
        return _sq(f, single)

    def density(self, xyz, t=0.0):
        x, single = prep_xyz(xyz, self.device)
        return _sq(self._density(x, t), single)

    def forceDeriv(self, xyz, t=0.0):
        x, single = prep_xyz(xyz, self.device)
        _, f, d = self._evaluate(x, False, True, True, t)
        return _sq(f, single), _sq(d, single)

    def evalDeriv(self, xyz, t=0.0):
        x, single = prep_xyz(xyz, self.device)
        p, f, d = self._evaluate(x, True, True, True, t)
        return _sq(p, single), _sq(f, single), _sq(d, single)

    def eval(self, xyz, pot: bool = False, acc: bool = False, der: bool = False, t=0.0):
        """Any combination of potential, acceleration (force) and force
        derivatives, in the order (phi, force, deriv); one launch serves all
        requested outputs (the reference launches twice for pot+acc,
        gpu_potential.py:928-932)."""
        if not (pot or acc or der):
            raise ValueError("eval(): at least one of pot, acc, der must be True.")
        x, single = prep_xyz(xyz, self.device)
        p, f, d = self._evaluate(x, pot, acc, der, t)
        res = [_sq(v, single) for v, w in ((p, pot), (f, acc), (d, der)) if w]
        return res[0] if len(res) == 1 else tuple(res)

    def eval_into(self, xyz: Tensor, phi: Optional[Tensor] = None, force: Optional[Tensor] = None,
                  deriv: Optional[Tensor] = None, t=0.0):
        """Evaluate into caller-owned device tensors (phi (N,), force (N,3),
        deriv (N,6), float64, contiguous, on this potential's device); only
        the given outputs are computed.  No allocation on this path."""
        if xyz.dtype != torch.float64 or xyz.dim() != 2 or xyz.shape[1] != 3 or \
                not xyz.is_contiguous() or xyz.device != self.device:
            raise ValueError("eval_into needs a contiguous float64 (N,3) tensor on the potential's device")
        n = xyz.shape[0]
        for name, tsr, shape in (("phi", phi, (n,)), ("force", force, (n, 3)), ("deriv", deriv, (n, 6))):
            if tsr is not None and (tuple(tsr.shape) != shape or tsr.dtype != torch.float64 or
                                    not tsr.is_contiguous() or tsr.device != self.device):
                raise ValueError(f"{name} must be a contiguous float64 tensor of shape {shape}")
        if phi is None and force is None and deriv is None:
            raise ValueError("eval_into(): no output given")
        return self._evaluate(xyz, phi is not None, force is not None, deriv is not None, t,
                              out=(phi, force, deriv))

    # ---- composition (gpu_potential.py:85-101) -------------------------------
    def __add__(self, other):
        from .composite import CompositePotentialGPU
        a = self._components if isinstance(self, CompositePotentialGPU) else [self]
        b = other._components if isinstance(other, CompositePotentialGPU) else [other]
        return CompositePotentialGPU(a + b)

    def __radd__(self, other):
        if isinstance(other, (int, float)) and other == 0:
            return self
        return NotImplemented


def alloc_outputs(n: int, device, want_phi: bool, want_force: bool, want_deriv: bool, out=None):
    """Output tensors for one evaluation; entries of ``out`` are reused."""
    kw = dict(dtype=torch.float64, device=device)
    o = out if out is not None else (None, None, None)
    phi = (o[0] if o[0] is not None else torch.empty(n, **kw)) if want_phi else None
    force = (o[1] if o[1] is not None else torch.empty((n, 3), **kw)) if want_force else None
    deriv = (o[2] if o[2] is not None else torch.empty((n, 6), **kw)) if want_deriv else None
    return phi, force, deriv


def norm_device(device) -> torch.device:
    """torch.device with an explicit index for GPUs ("cuda" -> "cuda:<current>"),
    so that device comparisons with tensors' devices hold."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def ptr(t: Optional[Tensor]):
    return None if t is None else t.data_ptr()
