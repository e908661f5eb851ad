"""The C-ABI library builds, loads and exports every entry point declared in
include/agama_gpu.h (no compute calls: runs without a GPU)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "agama_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(agx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "agx_multipole_eval" in syms and "agx_version" in syms


def test_library_exports_all_declared_symbols():
    from agama_gpu_amd import _lib
    path = _lib.library_path()
    if not path.exists():
        pytest.fail(f"{path} not built; run __graft_entry__.build()")
    L = C.CDLL(str(path))
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    for s in declared_symbols():
        assert s in _lib._SIGNATURES, f"{s} has no ctypes signature in _lib"


def test_version_and_error_strings():
    from agama_gpu_amd import _lib
    L = _lib.load()
    assert L.agx_version() == 10500
    assert b"NULL" in L.agx_error_string(-1)
    assert L.agx_error_string(12345) == b"unknown error code"


def test_argument_validation_without_gpu():
    """Invalid arguments are rejected on the host before any launch.  The
    positions are NULL in every call with N > 0: the pointer checks come after
    the descriptor and flag checks, so a check that regressed returns
    AGX_E_NULL (-1) instead of launching on a fake address."""
    from agama_gpu_amd import _lib
    L = _lib.load()
    d = _lib.agx_multipole()
    assert L.agx_multipole_eval(None, None, 0, None, None, None, 1, None) == -1
    assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 1, None) == -1
    d.poly_km = d.program = 8
    d.n_int, d.n_lm, d.lmax = 4, 1, 40
    assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 1, None) == -3
    d.lmax = 0
    assert L.agx_multipole_eval(C.byref(d), None, -1, None, None, None, 1, None) == -2
    assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 0, None) == -5
    assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 1, None) == -1
    assert L.agx_multipole_eval(C.byref(d), None, 0, None, None, None, 1, None) == 0
    for flags in (_lib.AGX_PROG_STRIDE1, _lib.AGX_PROG_COS_ONLY, _lib.AGX_PROG_STRIDE2):
        d.prog_flags = flags
        assert L.agx_multipole_eval(C.byref(d), None, 0, None, None, None, 1, None) == 0
    d.prog_flags = 8                    # no such program flag
    assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 1, None) == -5
    assert L.agx_multipole_density(C.byref(d), None, 10, 1.0, None, None) == -5
    d.prog_flags = 0
    assert L.agx_multipole_density(C.byref(d), None, 10, 1.0, None, None) == -1
    # a triaxial degree (bits 8-15) must be even and agree with the program's
    # size: degree 8 is 5 groups, 15 steps and 15 columns
    d.n_groups, d.n_steps, d.n_lm = 5, 15, 15
    for tl, ok in ((8, True), (7, False), (6, False), (10, False)):
        d.prog_flags = (tl << _lib.AGX_PROG_TRIAXIAL_SHIFT) | _lib.AGX_PROG_COS_ONLY | \
            _lib.AGX_PROG_STRIDE2
        assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 3, None) == \
            (-1 if ok else -5), tl
    d.n_lm = 14
    assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 3, None) == -5
    # a full degree (bits 16-23): degree 4 is 5 groups, 15 steps, 25 columns;
    # not together with a triaxial degree
    d.n_groups, d.n_steps, d.n_lm = 5, 15, 25
    for flags, ok in ((4 << 16, True), (5 << 16, False), ((4 << 16) | (8 << 8), False)):
        d.prog_flags = flags | _lib.AGX_PROG_STRIDE1
        assert L.agx_multipole_eval(C.byref(d), None, 10, None, None, None, 3, None) == \
            (-1 if ok else -5), hex(flags)
    d.prog_flags = 0


def test_component_flags_validation_without_gpu():
    """A Multipole component with an unknown program flag bit is rejected on
    the host before any launch (positions NULL, as above)."""
    from agama_gpu_amd import _lib
    L = _lib.load()
    comps = (_lib.agx_component * 1)()
    comps[0].type = _lib.AGX_T_MULTIPOLE
    comps[0].table = 8
    comps[0].ampl = comps[0].scale = 1.0
    comps[0].prog_flags = _lib.AGX_PROG_COS_ONLY | _lib.AGX_PROG_STRIDE2
    assert L.agx_composite_eval(comps, 1, None, 0, None, None, None, 1, None) == 0
    assert L.agx_composite_eval(comps, 1, None, 10, None, None, None, 1, None) == -1
    # the triaxial degree is accepted (the composite kernel does not use it)
    comps[0].prog_flags = (8 << 8) | _lib.AGX_PROG_COS_ONLY | _lib.AGX_PROG_STRIDE2
    assert L.agx_composite_eval(comps, 1, None, 10, None, None, None, 1, None) == -1
    comps[0].prog_flags = 8                # no such program flag
    assert L.agx_composite_eval(comps, 1, None, 10, None, None, None, 1, None) == -5


def test_descriptor_layout_matches_header():
    """The ctypes mirror of agx_multipole has the header's field order and
    C layout (no implicit padding before the doubles)."""
    from agama_gpu_amd import _lib
    text = (ROOT / "include" / "agama_gpu.h").read_text()
    body = re.search(r"typedef struct agx_multipole \{(.*?)\} agx_multipole;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b([A-Za-z_0-9]+)\s*[,;]", body)
    assert names == [f[0] for f in _lib.agx_multipole._fields_]
    assert _lib.agx_multipole.logr_min.offset == 56
    assert C.sizeof(_lib.agx_multipole) == 56 + 10 * 8 + 2 * 8


def test_adjoint_workspace_validation_without_gpu():
    """The Multipole adjoint needs its documented workspace (about 9 N bytes);
    positions NULL so that no failed check can launch."""
    from agama_gpu_amd import _lib
    L = _lib.load()
    n = 1000
    need = L.agx_multipole_work_size(n)
    assert need >= 9 * n and L.agx_multipole_work_size(-1) == 0
    d = _lib.agx_multipole()
    d.poly_km = d.program = 8
    d.n_int, d.n_lm, d.lmax = 4, 1, 0
    args = (C.byref(d), None, n, None, None, None, 8, 8)
    assert L.agx_multipole_adjoint(*args, None, need, 0, None) == -1
    assert L.agx_multipole_adjoint(*args, 8, need - 1, 0, None) == -7
    assert L.agx_multipole_adjoint(*args, 8, need, 0, None) == -1
    # flags: AGX_ADJ_CENTER_GRAD is known (the NULL positions then fail), other bits are not
    assert L.agx_multipole_adjoint(*args, 8, need, _lib.AGX_ADJ_CENTER_GRAD, None) == -1
    assert L.agx_multipole_adjoint(*args, 8, need, 4, None) == -5      # AGX_E_WHAT
    d.n_int = 128                       # n_int + 1 bins > 128
    assert L.agx_multipole_adjoint(*args, 8, need, 0, None) == -2


def test_noalloc_validation_without_gpu():
    """agx_multipole_eval_sorted_noalloc rejects N beyond its static scratch
    (AGX_E_SCRATCH) before its pointer checks, and NULL positions before any
    device call (both with N > 0 and NULL positions: nothing can launch)."""
    from agama_gpu_amd import _lib
    L = _lib.load()
    d = _lib.agx_multipole()
    d.poly_km = d.program = 8
    d.n_int, d.n_lm, d.lmax = 4, 1, 0
    f = L.agx_multipole_eval_sorted_noalloc
    assert f(C.byref(d), None, _lib.AGX_NOALLOC_MAX_N + 1, None, None, None, 1, None) == -8
    assert f(C.byref(d), None, _lib.AGX_NOALLOC_MAX_N, None, None, None, 1, None) == -1
    assert f(C.byref(d), None, 0, None, None, None, 1, None) == 0
    assert f(C.byref(d), None, 10, None, None, None, 0, None) == -5
    assert b"static scratch" in L.agx_error_string(-8)
