"""CylSplinePotentialGPU -- Agama CylSpline expansion on gfx950.

Mirrors the reference class (gpu_potential.py:1260-1432): construction from a
``CylSplineCoefs`` table or file, the Agama API, density from the Laplacian.
Tables are built once on the host (cylspline_prep.build_cylspline_data) and
held as device tensors; evaluation is one launch of the gfx950 kernel
(agx_cylspline_eval / agx_cylspline_density).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Union

import numpy as np
import torch

from . import _lib
from .base import PotentialBase, alloc_outputs, ptr
from .coefs import CylSplineCoefs, read_coefs
from .cylspline_prep import build_cylspline_data
from .leaves import Leaf


class CylSplinePotentialGPU(PotentialBase):
    """GPU evaluator for an Agama CylSpline expansion."""

    def __init__(self, coefs: CylSplineCoefs, device=None):
        super().__init__(device)
        d = build_cylspline_data(coefs)
        self._data = d
        dev = self.device
        self._lR = torch.as_tensor(d["lR_grid"], device=dev)
        self._lz = torch.as_tensor(d["lz_grid"], device=dev)
        self._nodes = torch.as_tensor(d["node_arr"].ravel(), device=dev)
        self._m = torch.as_tensor(d["m_arr"].astype(np.int32), device=dev)
        self._W = torch.as_tensor(d["W_outer"], device=dev)
        s = _lib.agx_cylspline()
        s.lR_grid, s.lz_grid = self._lR.data_ptr(), self._lz.data_ptr()
        s.nodes, s.m_arr, s.W_outer = self._nodes.data_ptr(), self._m.data_ptr(), self._W.data_ptr()
        s.nR, s.nz = len(d["lR_grid"]), len(d["lz_grid"])
        s.n_harm, s.log_scaling = d["n_harm"], d["log_scaling"]
        s.lmax_outer, s.mmax_outer = d["lmax_outer"], d["mmax_outer"]
        s.Rscale, s.r0_outer = d["Rscale"], d["r0_outer"]
        self._desc = s
        L = _lib.load()
        nbytes = int(L.agx_cylspline_pack_size())
        img = (C.c_uint8 * nbytes)()
        _lib.check(L.agx_cylspline_pack(C.byref(s), img, nbytes), "agx_cylspline_pack")
        self._dev_desc = torch.frombuffer(bytearray(img), dtype=torch.uint8).to(dev)

    @classmethod
    def from_file(cls, path: Union[str, Path], **kw) -> "CylSplinePotentialGPU":
        c = read_coefs(Path(path))
        if not isinstance(c, CylSplineCoefs):
            raise ValueError(f"{path} is not a CylSpline coefficient file")
        return cls(c, **kw)

    def __repr__(self):
        d = self._data
        return (f"CylSplinePotentialGPU(mmax={d['mmax']}, grid={len(d['lR_grid'])}x"
                f"{len(d['lz_grid'])}, Rscale={d['Rscale']:.4g})")

    def _leaves(self, t=0.0):
        return [Leaf(_lib.AGX_T_CYLSPLINE, table=self._dev_desc.data_ptr(), keepalive=self)]

    def _evaluate(self, xyz, want_phi, want_force, want_deriv, t=0.0, out=None):
        n = xyz.shape[0]
        phi, force, deriv = alloc_outputs(n, self.device, want_phi, want_force, want_deriv, out)
        what = (_lib.AGX_PHI if want_phi else 0) | (_lib.AGX_FORCE if want_force else 0) | \
               (_lib.AGX_DERIV if want_deriv else 0)
        rc = _lib.load().agx_cylspline_eval(C.byref(self._desc), xyz.data_ptr(), n, ptr(phi),
                                            ptr(force), ptr(deriv), what,
                                            _lib.stream_handle(self.device))
        _lib.check(rc, "agx_cylspline_eval")
        return phi, force, deriv

    def _density(self, xyz, t=0.0):
        rho = torch.empty(xyz.shape[0], dtype=torch.float64, device=self.device)
        rc = _lib.load().agx_cylspline_density(C.byref(self._desc), xyz.data_ptr(), xyz.shape[0],
                                               rho.data_ptr(), _lib.stream_handle(self.device))
        _lib.check(rc, "agx_cylspline_density")
        return rho
