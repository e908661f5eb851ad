"""Loader for the gfx950 C-ABI library (``libagama_gpu.so``).

The library is built in-tree by ``agama_gpu_amd/build.py`` (hipcc,
--offload-arch=gfx950).  There is no CPU fallback: if the library is missing
or no GPU is visible, evaluation raises.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np
import torch  # noqa: F401  -- loads the HIP runtime the library binds to

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libagama_gpu.so"

_lock = threading.Lock()
_lib = None


class AgxError(RuntimeError):
    """Error code returned by the native library."""


class agx_multipole(C.Structure):
    _fields_ = [
        ("poly_km", C.c_void_p), ("program", C.c_void_p),
        ("n_int", C.c_int32), ("n_lm", C.c_int32), ("n_groups", C.c_int32),
        ("n_steps", C.c_int32), ("lmax", C.c_int32),
        ("log_scaling", C.c_int32), ("inner_log", C.c_int32), ("outer_log", C.c_int32),
        ("prog_flags", C.c_uint32), ("pad", C.c_int32),
        ("logr_min", C.c_double), ("dlogr", C.c_double), ("inv_dlogr", C.c_double),
        ("invPhi0", C.c_double),
        ("inner_s", C.c_double), ("inner_U", C.c_double), ("inner_W", C.c_double),
        ("outer_s", C.c_double), ("outer_U", C.c_double), ("outer_W", C.c_double),
        ("law", C.c_void_p), ("origin", C.c_void_p),
    ]


class agx_multipole_source(C.Structure):
    _fields_ = [
        ("phi", C.c_void_p), ("dphi", C.c_void_p), ("r", C.c_void_p), ("cols", C.c_void_p),
        ("Mf", C.c_void_p), ("Md", C.c_void_p),
        ("nR", C.c_int32), ("n_cols", C.c_int32), ("n_lm", C.c_int32),
        ("log_scaling", C.c_int32), ("inner_mode", C.c_int32), ("inner_reset", C.c_int32),
        ("outer_mode", C.c_int32), ("pad", C.c_int32), ("dlogr", C.c_double),
    ]


class agx_cylspline(C.Structure):
    _fields_ = [
        ("lR_grid", C.c_void_p), ("lz_grid", C.c_void_p), ("nodes", C.c_void_p),
        ("m_arr", C.c_void_p), ("W_outer", C.c_void_p),
        ("nR", C.c_int32), ("nz", C.c_int32), ("n_harm", C.c_int32), ("log_scaling", C.c_int32),
        ("lmax_outer", C.c_int32), ("mmax_outer", C.c_int32),
        ("Rscale", C.c_double), ("r0_outer", C.c_double),
    ]


class agx_component(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("prog_flags", C.c_uint32), ("p", C.c_double * 8),
        ("table", C.c_void_p), ("ampl", C.c_double), ("scale", C.c_double),
        ("center", C.c_double * 3),
    ]


AGX_MAX_COMPONENTS = 16
(AGX_T_NFW, AGX_T_PLUMMER, AGX_T_HERNQUIST, AGX_T_DEHNEN, AGX_T_ISOCHRONE,
 AGX_T_MIYAMOTO_NAGAI, AGX_T_LOGHALO, AGX_T_DISK_ANSATZ, AGX_T_UNIFORM_ACCEL) = range(1, 10)
AGX_T_MULTIPOLE, AGX_T_CYLSPLINE = 16, 17

AGX_PHI, AGX_FORCE, AGX_DERIV = 1, 2, 4
AGX_WORK_SORTED = 1
AGX_ADJ_CENTER_GRAD = 2
AGX_PROG_STRIDE1 = 1
AGX_PROG_COS_ONLY = 2
AGX_PROG_STRIDE2 = 4
AGX_PROG_TRIAXIAL_SHIFT, AGX_PROG_TRIAXIAL_MASK = 8, 0xFF00
AGX_PROG_FULL_SHIFT, AGX_PROG_FULL_MASK = 16, 0xFF0000
AGX_OPT_MP_PATH, AGX_OPT_MP_TILE, AGX_OPT_MP_BIN_MIN_N, AGX_OPT_MP_BIN_MIN_TERMS = 1, 2, 3, 4
AGX_OPT_MP_LDS_GRID = 6
AGX_OPT_MP_BIN_GRID = 7
AGX_OPT_GRID_PER_CU = 8
AGX_OPT_SORT_GRID = 9
AGX_OPT_CYL_PATH = 10
AGX_OPT_ADJ_STAGE = 11
AGX_OPT_MP_TRIAXIAL = 12
AGX_BATCH_MAX_ITEM = 16384
AGX_NOALLOC_MAX_N = 11_000_000
AGX_E_SCRATCH = -8
AGX_BATCH_LANE, AGX_BATCH_LDS = 0, 1

# agx_sched_entry as a numpy record (int32 leaf, pad; ampl, scale, center[3])
SCHED_ENTRY_DTYPE = np.dtype([("leaf", "<i4"), ("pad", "<i4"), ("ampl", "<f8"), ("scale", "<f8"),
                              ("center", "<f8", (3,))])

# agx_batch_item as a numpy record (int64 start, int32 count, pot, path, pad)
BATCH_ITEM_DTYPE = np.dtype([("start", "<i8"), ("count", "<i4"), ("pot", "<i4"),
                             ("path", "<i4"), ("pad", "<i4")])


_options = {}


def set_option(key: int, value: int) -> None:
    """Process-wide tuning knob of the native library (see agx_set_option)."""
    check(load().agx_set_option(int(key), int(value)), "agx_set_option")
    _options[int(key)] = int(value)


def get_option(key: int, default: int = 0) -> int:
    """The value last set through set_option (the library's default otherwise)."""
    return _options.get(int(key), default)

# (name, restype, argtypes) of every exported entry point declared in
# include/agama_gpu.h
_SIGNATURES = {
    "agx_version": (C.c_int, []),
    "agx_set_option": (C.c_int, [C.c_int, C.c_int64]),
    "agx_error_string": (C.c_char_p, [C.c_int]),
    "agx_multipole_eval": (C.c_int, [C.POINTER(agx_multipole), C.c_void_p, C.c_int64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "agx_multipole_density": (C.c_int, [C.POINTER(agx_multipole), C.c_void_p, C.c_int64,
                                        C.c_double, C.c_void_p, C.c_void_p]),
    "agx_composite_eval": (C.c_int, [C.POINTER(agx_component), C.c_int, C.c_void_p, C.c_int64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                     C.c_void_p]),
    "agx_composite_density": (C.c_int, [C.POINTER(agx_component), C.c_int, C.c_void_p,
                                        C.c_int64, C.c_void_p, C.c_void_p]),
    "agx_multipole_adjoint": (C.c_int, [C.POINTER(agx_multipole), C.c_void_p, C.c_int64,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int64, C.c_uint32,
                                        C.c_void_p]),
    "agx_multipole_work_size": (C.c_int64, [C.c_int64]),
    "agx_multipole_eval_sorted": (C.c_int, [C.POINTER(agx_multipole), C.c_void_p, C.c_int64,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                            C.c_void_p, C.c_int64, C.c_void_p]),
    "agx_multipole_eval_sorted_noalloc": (C.c_int, [C.POINTER(agx_multipole), C.c_void_p,
                                                    C.c_int64, C.c_void_p, C.c_void_p,
                                                    C.c_void_p, C.c_uint32, C.c_void_p]),
    "agx_multipole_build": (C.c_int, [C.POINTER(agx_multipole_source), C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "agx_multipole_build_adjoint": (C.c_int, [C.POINTER(agx_multipole_source), C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p]),
    "agx_cylspline_eval": (C.c_int, [C.POINTER(agx_cylspline), C.c_void_p, C.c_int64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "agx_cylspline_density": (C.c_int, [C.POINTER(agx_cylspline), C.c_void_p, C.c_int64,
                                        C.c_void_p, C.c_void_p]),
    "agx_cylspline_adjoint": (C.c_int, [C.POINTER(agx_cylspline), C.c_void_p, C.c_int64,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "agx_cylspline_adjoint_work_size": (C.c_int64, [C.POINTER(agx_cylspline), C.c_int64]),
    "agx_cylspline_pack_size": (C.c_int64, []),
    "agx_cylspline_pack": (C.c_int, [C.POINTER(agx_cylspline), C.c_void_p, C.c_int64]),
    "agx_multipole_pack_size": (C.c_int64, []),
    "agx_multipole_pack": (C.c_int, [C.POINTER(agx_multipole), C.c_void_p, C.c_int64]),
    "agx_batch_plan": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "agx_batch_lds_table_doubles": (C.c_int64, []),
    "agx_leapfrog": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                               C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_int64, C.c_double, C.c_int32, C.c_void_p]),
    "agx_multipole_batch_eval": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                           C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_void_p]),
}


def library_path() -> Path:
    return LIB_PATH


def load():
    """Load (once, thread-safe) and return the ctypes handle."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the gfx950 library first "
                "(python -c 'import __graft_entry__ as g; g.build()' or "
                "python agama_gpu_amd/build.py)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "agx call") -> None:
    if rc != 0:
        msg = load().agx_error_string(rc).decode()
        raise AgxError(f"{what} failed: {msg} (code {rc})")


def stream_handle(device: torch.device):
    """hipStream_t of torch's current stream on ``device`` (as void*)."""
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("agama_gpu_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())
