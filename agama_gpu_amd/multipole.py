"""MultipolePotentialGPU -- Agama Multipole BFE on gfx950.

Mirrors the reference class ``MultipolePotentialGPU``
(gpu_potential.py:633-935): construction from a ``MultipoleCoefs`` table
(``prune_threshold`` kwarg, ``from_file``), and the Agama API.  Tables are
built once on the host (multipole_prep.build_multipole_data) and held as
torch tensors on the device; evaluation is one launch of the gfx950 kernel
through the C ABI (include/agama_gpu.h: agx_multipole_eval /
agx_multipole_density) on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Union

import numpy as np
import torch

from . import _lib
from .base import PotentialBase, alloc_outputs, ptr
from .coefs import MultipoleCoefs, read_coefs
from .multipole_prep import INV_4PIG, build_multipole_data, pack_program, program_flags


# Tables above this size do not fit the LDS-staged path; from this many points
# on they are evaluated through the interval-sorted path.
LDS_TABLE_BYTES = 160 * 1024 - 256
SORTED_MIN_N = 200_000


def eval_table(desc, data: dict, xyz, phi, force, deriv, what: int, device):
    """One evaluation of a Multipole descriptor: agx_multipole_eval, or for
    tables too large for LDS at large N (and whenever AGX_OPT_MP_PATH is 4)
    agx_multipole_eval_sorted with a workspace from torch's allocator.
    Returns that workspace (it holds the interval sort of xyz) or None."""
    n = xyz.shape[0]
    L = _lib.load()
    opt = _lib.get_option(_lib.AGX_OPT_MP_PATH)
    table = data["n_intervals"] * data["n_lm"] * 48
    sorted_ok = data["n_intervals"] + 1 <= 128 and n > 0
    use_sorted = sorted_ok and (opt == 4 or (opt == 0 and table > LDS_TABLE_BYTES and
                                               n >= SORTED_MIN_N))
    stream = _lib.stream_handle(device)
    if use_sorted:
        nwork = int(L.agx_multipole_work_size(n))
        work = torch.empty(nwork, dtype=torch.uint8, device=device)
        rc = L.agx_multipole_eval_sorted(C.byref(desc), xyz.data_ptr(), n, ptr(phi), ptr(force),
                                         ptr(deriv), what, work.data_ptr(), nwork, stream)
        _lib.check(rc, "agx_multipole_eval_sorted")
        return work
    rc = L.agx_multipole_eval(C.byref(desc), xyz.data_ptr(), n, ptr(phi), ptr(force),
                              ptr(deriv), what, stream)
    _lib.check(rc, "agx_multipole_eval")
    return None


class MultipolePotentialGPU(PotentialBase):
    """GPU evaluator for an Agama Multipole expansion (lmax <= 32)."""

    def __init__(self, coefs: MultipoleCoefs, prune_threshold: float = 1e-16, device=None):
        super().__init__(device)
        d = build_multipole_data(coefs, prune_threshold=prune_threshold)
        self._data = d
        prog = d["program"]
        dev = self.device
        self._poly = torch.as_tensor(d["poly_km"].ravel(), device=dev)
        self._program = torch.as_tensor(pack_program(prog).view(np.uint8), device=dev)
        s = _lib.agx_multipole()
        s.poly_km = self._poly.data_ptr()
        s.program = self._program.data_ptr()
        s.n_int = d["n_intervals"]
        s.n_lm = d["n_lm"]
        s.n_groups = len(prog.grp_m)
        s.n_steps = len(prog.step_l)
        s.prog_flags = program_flags(prog)
        s.lmax = d["lmax"]
        s.log_scaling = d["log_scaling"]
        s.inner_log = d["inner_log"]
        s.outer_log = d["outer_log"]
        for k in ("logr_min", "dlogr", "inv_dlogr", "invPhi0", "inner_s", "inner_U", "inner_W",
                  "outer_s", "outer_U", "outer_W"):
            setattr(s, k, d[k])
        self._desc = s
        # device image of the descriptor, for fused composites
        L = _lib.load()
        nbytes = int(L.agx_multipole_pack_size())
        img = (C.c_uint8 * nbytes)()
        _lib.check(L.agx_multipole_pack(C.byref(s), img, nbytes), "agx_multipole_pack")
        self._dev_desc = torch.frombuffer(bytearray(img), dtype=torch.uint8).to(dev)
        self._inv_4piG = INV_4PIG
        self.lmax = d["lmax"]
        self.n_lm = d["n_lm"]
        self.n_intervals = d["n_intervals"]

    # ---- constructors -------------------------------------------------------
    @classmethod
    def from_file(cls, path: Union[str, Path], **kw) -> "MultipolePotentialGPU":
        """Load from an Agama Multipole coefficient file (needs #dPhi/dr)."""
        c = read_coefs(Path(path))
        if not isinstance(c, MultipoleCoefs):
            raise ValueError(f"{path} is not a Multipole coefficient file")
        return cls(c, **kw)

    def __repr__(self):
        return f"MultipolePotentialGPU(lmax={self.lmax}, n_lm={self.n_lm}, n_int={self.n_intervals})"

    @property
    def descriptor(self):
        """The C-ABI descriptor (agx_multipole) of this potential's tables."""
        return self._desc

    # ---- primitives ---------------------------------------------------------
    def _leaves(self, t=0.0):
        from .leaves import Leaf
        return [Leaf(_lib.AGX_T_MULTIPOLE, table=self._dev_desc.data_ptr(), keepalive=self,
                     prog_flags=int(self._desc.prog_flags))]

    def _evaluate(self, xyz, want_phi, want_force, want_deriv, t=0.0, out=None):
        n = xyz.shape[0]
        phi, force, deriv = alloc_outputs(n, self.device, want_phi, want_force, want_deriv, out)
        what = (_lib.AGX_PHI if want_phi else 0) | (_lib.AGX_FORCE if want_force else 0) | \
               (_lib.AGX_DERIV if want_deriv else 0)
        eval_table(self._desc, self._data, xyz, phi, force, deriv, what, self.device)
        return phi, force, deriv

    def _density(self, xyz, t=0.0):
        n = xyz.shape[0]
        rho = torch.empty(n, dtype=torch.float64, device=self.device)
        rc = _lib.load().agx_multipole_density(C.byref(self._desc), xyz.data_ptr(), n,
                                               self._inv_4piG, rho.data_ptr(),
                                               _lib.stream_handle(self.device))
        _lib.check(rc, "agx_multipole_density")
        return rho
