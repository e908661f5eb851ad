"""ORACLE -- test infrastructure only (ctypes front end of mp_oracle.c).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product package.  Evaluates the C restatement of the
reference's Multipole path (gpu_potential.py:183-626,
_multipole_potential_kernel.cu:100-958) on the CPU with OpenMP.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libmp_oracle.so"

G_AGAMA = 4.30091727067736e-06
INV_4PIG = 1.0 / (4.0 * np.pi * G_AGAMA)


class _OrcMP(C.Structure):
    _fields_ = [
        ("n_int", C.c_int), ("n_lm", C.c_int), ("lmax", C.c_int), ("log_scaling", C.c_int),
        ("inner_log", C.c_int), ("outer_log", C.c_int),
        ("logr_min", C.c_double), ("dlogr", C.c_double), ("inv_dlogr", C.c_double),
        ("invPhi0", C.c_double),
        ("inner_s", C.c_double), ("inner_U", C.c_double), ("inner_W", C.c_double),
        ("outer_s", C.c_double), ("outer_U", C.c_double), ("outer_W", C.c_double),
        ("lm_l", C.POINTER(C.c_int)), ("lm_m", C.POINTER(C.c_int)),
        ("poly", C.POINTER(C.c_double)),
    ]


_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "mp_oracle.c").stat().st_mtime:
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.orc_mp_build.restype = C.c_int
        L.orc_mp_build.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                   C.c_void_p, C.c_void_p, C.c_double, C.POINTER(_OrcMP)]
        L.orc_mp_free.argtypes = [C.POINTER(_OrcMP)]
        L.orc_mp_eval.argtypes = [C.POINTER(_OrcMP), C.c_void_p, C.c_int64, C.c_int,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int]
        L.orc_mp_sizeof.restype = C.c_int
        L.orc_prefact.restype = C.c_double
        L.orc_prefact.argtypes = [C.c_int]
        L.orc_coef.restype = C.c_double
        L.orc_coef.argtypes = [C.c_int]
        assert L.orc_mp_sizeof() == C.sizeof(_OrcMP), "oracle struct layout mismatch"
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class OracleMultipole:
    """CPU restatement of the reference Multipole potential from raw tables."""

    def __init__(self, r, phi, dphi_dr, lm_labels, prune_threshold: float = 1e-16):
        L = lib()
        r = np.ascontiguousarray(r, np.float64)
        phi = np.ascontiguousarray(phi, np.float64)
        dphi = np.ascontiguousarray(dphi_dr, np.float64)
        lo = np.ascontiguousarray([l for l, _ in lm_labels], np.int32)
        mo = np.ascontiguousarray([m for _, m in lm_labels], np.int32)
        self._t = _OrcMP()
        rc = L.orc_mp_build(_ptr(r), _ptr(phi), _ptr(dphi), len(r), phi.shape[1],
                            _ptr(lo), _ptr(mo), float(prune_threshold), C.byref(self._t))
        if rc != 0:
            raise ValueError(f"oracle table build failed (code {rc})")
        t = self._t
        self.n_int, self.n_lm = t.n_int, t.n_lm
        self.poly = np.ctypeslib.as_array(t.poly, shape=(t.n_lm * t.n_int * 6,)).reshape(t.n_lm, t.n_int, 6).copy()
        self.lm_l = np.ctypeslib.as_array(t.lm_l, shape=(t.n_lm,)).copy()
        self.lm_m = np.ctypeslib.as_array(t.lm_m, shape=(t.n_lm,)).copy()
        self.scalars = {k: getattr(t, k) for k, _ in _OrcMP._fields_
                        if k not in ("lm_l", "lm_m", "poly")}

    @classmethod
    def from_coefs(cls, coefs, prune_threshold: float = 1e-16):
        return cls(coefs.R_grid, coefs.phi, coefs.dphi_dr, coefs.lm_labels, prune_threshold)

    def __del__(self):
        try:
            lib().orc_mp_free(C.byref(self._t))
        except Exception:
            pass

    def _run(self, xyz, mode: int, threads: int = 0):
        xyz = np.ascontiguousarray(np.asarray(xyz, np.float64).reshape(-1, 3))
        n = xyz.shape[0]
        phi = np.empty(n)
        grad = np.empty((n, 3)) if mode in (1, 2) else None
        hess = np.empty((n, 6)) if mode == 2 else None
        lib().orc_mp_eval(C.byref(self._t), _ptr(xyz), n, mode, _ptr(phi),
                          _ptr(grad) if grad is not None else None,
                          _ptr(hess) if hess is not None else None, INV_4PIG, int(threads))
        return phi, grad, hess

    # Agama-convention outputs: force = -grad Phi, deriv = -Hessian
    def potential(self, xyz, threads: int = 0):
        return self._run(xyz, 0, threads)[0]

    def pot_force(self, xyz, threads: int = 0):
        p, g, _ = self._run(xyz, 1, threads)
        return p, -g

    def eval_deriv(self, xyz, threads: int = 0):
        p, g, h = self._run(xyz, 2, threads)
        return p, -g, -h

    def density(self, xyz, threads: int = 0):
        return self._run(xyz, 3, threads)[0]
