"""Ragged ensembles of independent Multipole potentials in one launch.

``MultipoleEnsembleGPU(potentials, counts)`` evaluates potential p at its own
``counts[p]`` rows of one concatenated (sum counts, 3) position array -- every
potential with its own lmax, radial grid and particle count -- with one
persistent launch of the gfx950 batch kernel (agx_multipole_batch_eval).  The
host side packs the potentials' descriptor images into one device buffer and
plans work items (agx_batch_plan: rows per item and evaluation path per
potential, items ordered longest first); the device takes items from a queue.

``eval_separately`` is the per-potential form -- one plain launch per
potential on its slice -- which the batched form replaces.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .base import ptr


def _cost_per_particle(pot) -> float:
    """Interior FP64 flops of pot+force for one particle (80 + 40 n + 12 S + 5 G)."""
    prog = pot._data["program"]
    G = len(prog.grp_m)
    S = int(sum(int(prog.step_l[prog.grp_step0[g + 1] - 1]) - int(prog.grp_m[g]) for g in range(G)))
    return 80.0 + 40.0 * pot.n_lm + 12.0 * S + 5.0 * G


class MultipoleEnsembleGPU:
    """Many Multipole potentials, each on its own rows, in one launch.

    potentials: MultipolePotentialGPU objects on one device (their tables stay
    owned by them); counts: rows per potential, in order; choices: optional
    (path, rows per item) per potential (default: ``autotune.default_choice``;
    ``tune=True`` measures them with ``autotune``)."""

    def __init__(self, potentials: Sequence, counts: Sequence[int],
                 choices: Optional[Sequence[Tuple[int, int]]] = None, tune: bool = False):
        from . import autotune
        if len(potentials) != len(counts):
            raise ValueError("one particle count per potential is required")
        if len(potentials) == 0:
            raise ValueError("empty ensemble")
        self.potentials = list(potentials)
        self.device = self.potentials[0].device
        for p in self.potentials:
            if p.device != self.device:
                raise ValueError("all potentials of an ensemble must live on one device")
        counts = np.asarray(counts, dtype=np.int64)
        if np.any(counts < 0):
            raise ValueError("particle counts must be >= 0")
        self.counts = counts
        self.offsets = np.concatenate([[0], np.cumsum(counts)])
        self.n_rows = int(self.offsets[-1])
        if choices is None:
            choices = autotune.tune_ensemble(self.potentials) if tune else \
                [autotune.default_choice(p) for p in self.potentials]
        self.choices = [(int(a), int(b)) for a, b in choices]
        L = _lib.load()
        n_pot = len(self.potentials)
        size = int(L.agx_multipole_pack_size())
        self._stride = (size + 15) // 16 * 16
        img = np.zeros((n_pot, self._stride), np.uint8)
        for j, p in enumerate(self.potentials):
            buf = (C.c_uint8 * size)()
            _lib.check(L.agx_multipole_pack(C.byref(p.descriptor), buf, size), "agx_multipole_pack")
            img[j, :size] = np.frombuffer(buf, np.uint8)
        self._desc = torch.as_tensor(img.ravel(), device=self.device)
        # the program flags every potential satisfies (selects the kernel variant)
        self._prog_flags = 0xFFFFFFFF
        for p in self.potentials:
            self._prog_flags &= int(p.descriptor.prog_flags)
        # (the triaxial degrees, which differ between potentials, go with
        # each potential's items instead: bits 8-15 of an LDS item's path)
        self._prog_flags &= ~(_lib.AGX_PROG_TRIAXIAL_MASK | _lib.AGX_PROG_FULL_MASK) & 0xFFFFFFFF
        path = np.array([c[0] for c in self.choices], np.int32)
        for j, p in enumerate(self.potentials):
            if path[j] == _lib.AGX_BATCH_LDS:
                path[j] |= int(p.descriptor.prog_flags) & _lib.AGX_PROG_TRIAXIAL_MASK
        chunk = np.array([c[1] for c in self.choices], np.int32)
        cost = np.array([_cost_per_particle(p) for p in self.potentials], np.float64)
        n_items = C.c_int64(0)
        _lib.check(L.agx_batch_plan(n_pot, counts.ctypes.data, chunk.ctypes.data, path.ctypes.data,
                                    cost.ctypes.data, None, 0, C.byref(n_items)), "agx_batch_plan")
        items = np.zeros(n_items.value, _lib.BATCH_ITEM_DTYPE)
        _lib.check(L.agx_batch_plan(n_pot, counts.ctypes.data, chunk.ctypes.data, path.ctypes.data,
                                    cost.ctypes.data, items.ctypes.data if len(items) else None,
                                    len(items), C.byref(n_items)), "agx_batch_plan")
        self.items = items
        self._items = torch.as_tensor(items.view(np.uint8).copy(), device=self.device)
        self._queue = torch.zeros(1, dtype=torch.int32, device=self.device)

    def __len__(self):
        return len(self.potentials)

    def slice(self, p: int) -> slice:
        return slice(int(self.offsets[p]), int(self.offsets[p + 1]))

    def _check_rows(self, xyz):
        if not (isinstance(xyz, torch.Tensor) and xyz.device == self.device and
                xyz.dtype == torch.float64 and xyz.shape == (self.n_rows, 3) and xyz.is_contiguous()):
            raise ValueError(f"xyz must be a contiguous float64 ({self.n_rows}, 3) tensor on "
                             f"{self.device}")

    def eval_into(self, xyz, phi=None, force=None, deriv=None):
        """One launch over every potential's rows; writes the given outputs."""
        self._check_rows(xyz)
        what = (_lib.AGX_PHI if phi is not None else 0) | \
               (_lib.AGX_FORCE if force is not None else 0) | \
               (_lib.AGX_DERIV if deriv is not None else 0)
        rc = _lib.load().agx_multipole_batch_eval(
            self._desc.data_ptr(), self._stride, len(self.potentials), self._items.data_ptr(),
            len(self.items), xyz.data_ptr(), self.n_rows, ptr(phi), ptr(force), ptr(deriv), what,
            self._prog_flags, self._queue.data_ptr(), _lib.stream_handle(self.device))
        _lib.check(rc, "agx_multipole_batch_eval")
        return phi, force, deriv

    def eval(self, xyz, pot: bool = True, acc: bool = True, der: bool = False):
        """(phi, force, deriv) over the concatenated rows (None where not asked)."""
        xyz = torch.as_tensor(xyz, dtype=torch.float64, device=self.device).contiguous()
        n = self.n_rows
        phi = torch.empty(n, dtype=torch.float64, device=self.device) if pot else None
        force = torch.empty((n, 3), dtype=torch.float64, device=self.device) if acc else None
        deriv = torch.empty((n, 6), dtype=torch.float64, device=self.device) if der else None
        return self.eval_into(xyz, phi, force, deriv)

    def eval_separately(self, xyz, phi=None, force=None, deriv=None):
        """The same result by one plain launch per potential on its slice."""
        self._check_rows(xyz)
        for j, p in enumerate(self.potentials):
            sl = self.slice(j)
            if sl.stop == sl.start:
                continue
            p.eval_into(xyz[sl], phi=None if phi is None else phi[sl],
                        force=None if force is None else force[sl],
                        deriv=None if deriv is None else deriv[sl])
        return phi, force, deriv
