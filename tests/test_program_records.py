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


@pytest.mark.parametrize("lmax,sym", [(8, "triaxial"), (8, "none"), (32, "none")])
def test_records_encode_program(lmax, sym):
    from agama_gpu_amd.multipole_prep import PROGRAM_RECORD, pack_program
    l, m, prog = _program(lmax, sym)
    rec = pack_program(prog)
    assert rec.dtype == PROGRAM_RECORD and rec.dtype.itemsize == 32
    G, S = len(prog.grp_m), len(prog.step_l)
    assert len(rec) == G + S + 2
    assert not any(rec[G].tolist()) and not any(rec[-1].tolist())
    seen = set()
    for g in range(G):
        mm, j0, j1 = rec[g]["i0"], rec[g]["i1"], rec[g]["i2"]
        assert rec[g]["d0"] == prog.qmm[g]
        for j in range(j0, j1):
            s = rec[G + 1 + j]
            assert s["i0"] == mm + (j - j0)
            for off, sign in ((s["i1"], 1), (s["i2"], -1)):
                if off < 0:
                    continue
                assert off % 6 == 0
                c = off // 6
                assert (l[c], m[c]) == (s["i0"], sign * mm if sign > 0 else -mm)
                seen.add(c)
    assert seen == set(range(1, len(l)))          # every non-monopole term once


def test_records_reproduce_legendre():
    """Run the recurrence from the records at a few cos(theta) and compare
    mul_m x normalised P_l^m (Condon-Shortley) from scipy."""
    from agama_gpu_amd.multipole_prep import pack_program
    _, _, prog = _program(12, "none")
    rec = pack_program(prog)
    G = len(prog.grp_m)
    for ct in (-0.93, -0.2, 0.0, 0.41, 0.99):
        st = math.sqrt(1 - ct * ct)
        for g in range(G):
            mm, j0, j1, q = rec[g]["i0"], rec[g]["i1"], rec[g]["i2"], rec[g]["d0"]
            Q1, Q2 = q * st ** mm, 0.0
            for j in range(j0, j1):
                s = rec[G + 1 + j]
                ll = s["i0"]
                if ll > mm:
                    Q1, Q2 = s["d0"] * ct * Q1 - s["d1"] * Q2, Q1
                norm = math.sqrt((2 * ll + 1) / (4 * math.pi) * math.factorial(ll - mm) /
                                 math.factorial(ll + mm))
                mul = 2 * math.sqrt(math.pi) if mm == 0 else 2 * math.sqrt(2 * math.pi)
                want = mul * norm * lpmv(mm, ll, ct)
                assert abs(Q1 - want) <= 1e-12 * max(1.0, abs(want)), (ll, mm, ct)
