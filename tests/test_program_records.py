"""Layout of the Multipole kernels' Legendre program records
(multipole_prep.pack_program, documented at agx_multipole.program in
include/agama_gpu.h), checked field by field against the program arrays and
against the recurrence they encode: running the records reproduces
scipy's associated Legendre functions in Agama's normalisation."""
import math

import numpy as np
import pytest
from scipy.special import lpmv


def _program(lmax, sym="none"):
    from agama_gpu_amd.expansion import lm_labels
    from agama_gpu_amd.multipole_prep import _sort_key_order, build_legendre_program
    lab = [(l, m) for l, m in lm_labels(lmax)
           if sym == "none" or (l % 2 == 0 and m % 2 == 0 and m >= 0)]
    l = np.array([a for a, _ in lab], np.int32)
    m = np.array([b for _, b in lab], np.int32)
    o = _sort_key_order(l, m)
    return l[o], m[o], build_legendre_program(l[o], m[o])


@pytest.mark.parametrize("lmax,sym", [(8, "triaxial"), (8, "none"), (32, "none"),
                                      (32, "triaxial")])
def test_records_encode_program(lmax, sym):
    from agama_gpu_amd.multipole_prep import PROGRAM_RECORD, pack_program, unpack_offsets
    l, m, prog = _program(lmax, sym)
    rec = pack_program(prog)
    assert rec.dtype == PROGRAM_RECORD and rec.dtype.itemsize == 32
    G, S = len(prog.grp_m), len(prog.step_l)
    assert len(rec) == G + S + 2
    assert not any(rec[G].tolist()) and not any(rec[-1].tolist())
    seen = set()
    for g in range(G):
        mm, j0, j1, stride = rec[g]["i0"], rec[g]["i1"], rec[g]["i2"], rec[g]["i3"]
        # stride 2 exactly when every term of the group has l - m even
        ls = [int(a) for a, b in zip(l[1:], m[1:]) if abs(int(b)) == mm]
        assert stride == (2 if all((a - mm) % 2 == 0 for a in ls) else 1)
        assert rec[g]["d0"] == prog.qmm[g] and rec[g]["d1"] == float(mm)
        for j in range(j0, j1):
            s = rec[G + 1 + j]
            assert s["i0"] == mm + stride * (j - j0)
            oc, os_ = unpack_offsets(np.array([s["i1"]]))
            for off, sign in ((int(oc[0]), 1), (int(os_[0]), -1)):
                if off < 0:
                    continue
                assert off % 6 == 0
                c = off // 6
                assert (l[c], m[c]) == (s["i0"], sign * mm if sign > 0 else -mm)
                seen.add(c)
    assert seen == set(range(1, len(l)))          # every non-monopole term once


@pytest.mark.parametrize("lmax,sym", [(12, "none"), (12, "triaxial"), (32, "triaxial")])
def test_records_reproduce_legendre(lmax, sym):
    """Run the recurrence from the records at a few theta and compare
    mul_m x normalised P_l^m (Condon-Shortley) from scipy, and its theta
    derivative with a central difference of the same.  Every step is
    P_l = alpha (c + beta) P_{l-s} + gamma P_{l-2s} with c = cos theta for
    stride s = 1 (where beta = 0) and cos^2 theta for the two-step
    (same-parity) groups, differentiated as dP_l = alpha (c + beta) dP_{l-s}
    + alpha c' P_{l-s} + gamma dP_{l-2s}."""
    from agama_gpu_amd.multipole_prep import pack_program
    _, _, prog = _program(lmax, sym)
    rec = pack_program(prog)
    G = len(prog.grp_m)

    def want(ll, mm, th):
        norm = math.sqrt((2 * ll + 1) / (4 * math.pi) * math.factorial(ll - mm) /
                         math.factorial(ll + mm))
        mul = 2 * math.sqrt(math.pi) if mm == 0 else 2 * math.sqrt(2 * math.pi)
        return mul * norm * lpmv(mm, ll, math.cos(th))

    for th in (2.76, 1.77, 0.5 * math.pi, 1.15, 0.14):
        ct, st = math.cos(th), math.sin(th)
        for g in range(G):
            mm, j0, j1, q = rec[g]["i0"], rec[g]["i1"], rec[g]["i2"], rec[g]["d0"]
            two = rec[g]["i3"] == 2
            c, dc = (ct * ct, -2 * ct * st) if two else (ct, -st)
            Q1, Q2 = q * st ** mm, 0.0
            dQ1, dQ2 = q * mm * st ** (mm - 1) * ct if mm else 0.0, 0.0
            for j in range(j0, j1):
                s = rec[G + 1 + j]
                ll = s["i0"]
                if ll > mm:
                    assert two or s["d1"] == 0.0
                    t = s["d0"] * (c + s["d1"])
                    dQ1, dQ2 = t * dQ1 + s["d0"] * dc * Q1 + s["g"] * dQ2, dQ1
                    Q1, Q2 = t * Q1 + s["g"] * Q2, Q1
                w = want(ll, mm, th)
                assert abs(Q1 - w) <= 1e-12 * max(1.0, abs(w)), (ll, mm, th)
                h = 1e-5
                dw = (want(ll, mm, th + h) - want(ll, mm, th - h)) / (2 * h)
                assert abs(dQ1 - dw) <= 1e-7 * max(1.0, abs(dw)), (ll, mm, th, dQ1, dw)


@pytest.mark.parametrize("lmax", [2, 4, 6, 8, 10, 12, 32])
def test_triaxial_degree_of_triaxial_programs(lmax):
    """Every even-even m >= 0 harmonic up to an even lmax: the triaxial
    program of that degree, whose group g starts at step tri_step0(L, g) =
    g (L/2 + 1) - g (g - 1) / 2 (the kernels' compile-time layout), with the
    cos term of step j in column j; in program_flags bits 8-15."""
    from agama_gpu_amd.multipole_prep import program_flags, triaxial_degree
    l, m, prog = _program(lmax, "triaxial")
    assert triaxial_degree(prog) == lmax
    assert program_flags(prog) >> 8 == lmax and program_flags(prog) & 0xFF == 2 | 4
    G = lmax // 2 + 1
    assert list(prog.grp_step0) == [g * (lmax // 2 + 1) - g * (g - 1) // 2 for g in range(G + 1)]
    assert list(prog.step_cos) == [-1] + list(range(1, len(prog.step_l)))
    # an odd lmax adds nothing that triaxial symmetry allows
    if lmax < 32:
        assert triaxial_degree(_program(lmax + 1, "triaxial")[2]) == lmax


def test_triaxial_degree_rejects_other_programs():
    """A dense program, one with a term missing (the kernels would read the
    wrong column), one with a sin term, one with an odd-l term, the monopole
    alone: no triaxial degree."""
    from agama_gpu_amd.expansion import lm_labels
    from agama_gpu_amd.multipole_prep import (_sort_key_order, build_legendre_program,
                                              triaxial_degree)

    def prog_of(keep, lmax=8):
        lab = [(a, b) for a, b in lm_labels(lmax) if keep(a, b)]
        l = np.array([a for a, _ in lab], np.int32)
        m = np.array([b for _, b in lab], np.int32)
        o = _sort_key_order(l, m)
        return build_legendre_program(l[o], m[o])

    tri = lambda a, b: a % 2 == 0 and b % 2 == 0 and b >= 0          # noqa: E731
    assert triaxial_degree(prog_of(tri)) == 8
    assert triaxial_degree(prog_of(lambda a, b: True)) == 0
    assert triaxial_degree(prog_of(lambda a, b: tri(a, b) and (a, b) != (6, 4))) == 0
    assert triaxial_degree(prog_of(lambda a, b: tri(a, b) and (a, b) != (8, 8))) == 0
    assert triaxial_degree(prog_of(lambda a, b: tri(a, b) or (a, b) == (4, -2))) == 0
    assert triaxial_degree(prog_of(lambda a, b: tri(a, b) or (a, b) == (3, 0))) == 0
    assert triaxial_degree(prog_of(lambda a, b: a == 0)) == 0


@pytest.mark.parametrize("lmax", [1, 2, 5, 8, 10, 32])
def test_full_degree_of_full_programs(lmax):
    """Every harmonic up to lmax: the full program of that degree, whose
    group g starts at step full_step0(L, g) = g (L + 1) - g (g - 1) / 2, with
    the cos term of step j in column j (j <= L) or 2j - L - 1 and the sin term
    in column 2j - L; in program_flags bits 16-23, and no triaxial degree."""
    from agama_gpu_amd.multipole_prep import full_degree, program_flags, triaxial_degree
    l, m, prog = _program(lmax, "none")
    assert full_degree(prog) == lmax and triaxial_degree(prog) == 0
    assert program_flags(prog) >> 16 == lmax and (program_flags(prog) >> 8) & 0xFF == 0
    G = lmax + 1
    assert list(prog.grp_step0) == [g * (lmax + 1) - g * (g - 1) // 2 for g in range(G + 1)]


def test_full_degree_rejects_other_programs():
    """A triaxial program, a full program with one term missing or a sin term
    missing, the monopole alone: no full degree."""
    from agama_gpu_amd.expansion import lm_labels
    from agama_gpu_amd.multipole_prep import _sort_key_order, build_legendre_program, full_degree

    def prog_of(keep, lmax=6):
        lab = [(a, b) for a, b in lm_labels(lmax) if keep(a, b)]
        l = np.array([a for a, _ in lab], np.int32)
        m = np.array([b for _, b in lab], np.int32)
        o = _sort_key_order(l, m)
        return build_legendre_program(l[o], m[o])

    assert full_degree(prog_of(lambda a, b: True)) == 6
    assert full_degree(prog_of(lambda a, b: a % 2 == 0 and b % 2 == 0 and b >= 0)) == 0
    assert full_degree(prog_of(lambda a, b: (a, b) != (4, 2))) == 0
    assert full_degree(prog_of(lambda a, b: (a, b) != (3, -1))) == 0
    assert full_degree(prog_of(lambda a, b: (a, b) != (6, 6))) == 0
    assert full_degree(prog_of(lambda a, b: a == 0)) == 0
