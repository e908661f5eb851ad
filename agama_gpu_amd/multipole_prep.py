"""Host-side Multipole table builder (CPU, once per potential).

Restates the reference's ``_build_multipole_data`` (gpu_potential.py:408-626)
and its helpers ``_compute_invPhi0`` (183-263), ``_compute_outer_extrap``
(266-328) and ``_solve_quintic_d2`` (331-401): log-grid check and resample,
outer/inner l=0 extrapolation coefficients, (|m|, l, cos<sin) ordering,
zero-column pruning, Agama log-scaling, the quintic C2 spline and its
per-interval Horner coefficients.

On top of the reference tables it emits the layout the gfx950 kernels read:

* ``poly_km`` -- interval-major copy of the Horner table, shape
  (n_int, n_lm, 6): one radial interval is one contiguous slab, so a
  wavefront whose particles share an interval reads it with scalar loads.
* a Legendre "program": the active harmonics grouped by |m|, each group a
  run of recurrence steps l = |m| .. l_last with the fully normalised
  recurrence constants (alpha_l, beta_l) and the table column of the cos
  and sin term active at that l (or -1).  Agama's angular factor
  ``mul * PREFACT * raw P_l^m`` (reference kernel
  _multipole_potential_kernel.cu:54-88, 300-363) is folded into the seed
  ``qmm[m]`` so the kernel never forms factorials, square roots or the
  per-step normalisation update the reference recomputes per particle.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np

LMAX_LIMIT = 32

# kpc (km/s)^2 / Msol -- reference gpu_potential.py:157
G_AGAMA = 4.30091727067736e-06
INV_4PIG = 1.0 / (4.0 * math.pi * G_AGAMA)

_EPS = np.finfo(np.float64).eps


def inner_extrapolation(phi0: np.ndarray, dphi0_dr: np.ndarray, r: np.ndarray
                        ) -> Tuple[float, float, float, float, bool]:
    """(invPhi0, s, U, W, log_mode) of the inner l=0 law (v = 0).

    Power-law mode: Phi = U (r/r0)^s + W.  Log mode (s == v == 0):
    Phi = U ln(r/r0) + W.

    Follows gpu_potential.py:183-263.  The reference computes
    s = v + (A - W_k(A e^A)) / lnr with the Lambert-W branch k that contains A
    (``0 if A > -1 else -1``, gpu_potential.py:226-229).  On that branch
    W_k(A e^A) = A identically, so s = v exactly; the reference's
    floating-point evaluation leaves s = v + O(1e-16) noise and then divides
    by (s - v).  The exact-arithmetic value of its formulas is the s == v
    model above (U = r1 dPhi1 - v Phi1, W = Phi1, gpu_potential.py:240-242),
    which is what this function returns.  The s=2 constant-density-core
    override (245-251) and the invPhi0 safety reset (254-261) follow.
    """
    P1, P2 = float(phi0[0]), float(phi0[1])
    d1, d2 = float(dphi0_dr[0]), float(dphi0_dr[1])
    r1, r2 = float(r[0]), float(r[1])
    lnr = math.log(r2 / r1)
    num1 = r1 * d1
    safety = 100.0 * _EPS
    if num1 == 0.0 or abs(P1 - P2) < safety * max(abs(P1), abs(P2)):
        return 0.0, 0.0, 0.0, P1, False
    A = lnr * num1 / (P1 - P2)
    if not math.isfinite(A) or A >= 0.0:
        return 0.0, 0.0, 0.0, P1, False
    # trivial root: s = v = 0, log mode
    s, U, W, log_mode = 0.0, num1, P1, True
    # power-law (here: log-law) vs cubic prediction of dPhi/dr at node 2
    pred_pow = num1 / r2
    pred_cub = (r2 / r1 * (6.0 * r1 * (P2 - P1) / (r2 - r1) - d1 * (2 * r1 + r2))) / (2 * r2 + r1)
    if abs(d2 - pred_cub) < abs(d2 - pred_pow):
        s, U, log_mode = 2.0, 0.5 * num1, False
        W = P1 - U
    if s <= 0.0:
        return 0.0, float(s), float(U), float(W), log_mode
    inv = 1.0 / W
    if np.any(phi0 * inv >= 1.0):
        inv = 0.0
    return float(inv), float(s), float(U), float(W), log_mode


def outer_extrapolation(phi0: np.ndarray, dphi0_dr: np.ndarray, r: np.ndarray
                        ) -> Tuple[float, float, float, bool]:
    """(s, U, W, log_mode) of the outer l=0 law (v = -1).

    Power-law mode: Phi = W (r/rN)^-1 + U (r/rN)^s.  Log mode (s == v):
    Phi = (W + U ln(r/rN)) (r/rN)^-1.

    Follows gpu_potential.py:266-328 (round-off fallback to a pure Keplerian
    tail).  As for the inner law, the Lambert-W branch the reference selects
    (gpu_potential.py:315-316) yields s = v exactly; its s == v coefficients
    (322-324) with the log term are the exact value of its formulas.
    """
    v = -1.0
    P1, P2 = float(phi0[-1]), float(phi0[-2])
    d1 = float(dphi0_dr[-1])
    r1, r2 = float(r[-1]), float(r[-2])
    lnr = math.log(r2 / r1)
    num1, num2 = r1 * d1, v * P1
    den1, den2 = P1, P2 * math.exp(-v * lnr)
    safety = 100.0 * _EPS
    roundoff = (abs(num1 - num2) < max(abs(num1), abs(num2)) * safety or
                abs(den1 - den2) < max(abs(den1), abs(den2)) * safety)
    if not math.isfinite(P1) or roundoff or den1 == den2:
        return 0.0, 0.0, P1, False
    A = lnr * (num1 - num2) / (den1 - den2)
    if not math.isfinite(A) or A >= 0.0:
        return 0.0, 0.0, P1, False
    return v, num1 - v * P1, P1, True


def quintic_second_derivs(x: np.ndarray, f: np.ndarray, df: np.ndarray) -> np.ndarray:
    """Second derivatives of the C2 quintic spline through (f, f') with
    f'''' = 0 at both ends (gpu_potential.py:331-401).  f, df: (n, k)."""
    from scipy.linalg import solve_banded
    n = len(x)
    diag = np.zeros(n)
    up = np.zeros(n)     # A[i, i+1]
    lo = np.zeros(n)     # A[i, i-1]
    rhs = np.zeros_like(f)
    for i in range(n):
        if i > 0:
            h = 1.0 / (x[i] - x[i - 1])
            diag[i] += 3.0 * h
            lo[i] = -h
            rhs[i] -= (20.0 * (f[i] - f[i - 1]) * h - 12.0 * df[i] - 8.0 * df[i - 1]) * h * h
        if i < n - 1:
            h = 1.0 / (x[i + 1] - x[i])
            diag[i] += 3.0 * h
            up[i] = -h
            rhs[i] += (20.0 * (f[i + 1] - f[i]) * h - 12.0 * df[i] - 8.0 * df[i + 1]) * h * h
    h0 = 1.0 / (x[1] - x[0])
    up[0] = -2.0 * h0
    rhs[0] = (30.0 * (f[1] - f[0]) * h0 - 14.0 * df[1] - 16.0 * df[0]) * h0 * h0
    hn = 1.0 / (x[n - 1] - x[n - 2])
    lo[n - 1] = -2.0 * hn
    rhs[n - 1] = (-30.0 * (f[n - 1] - f[n - 2]) * hn + 14.0 * df[n - 2] + 16.0 * df[n - 1]) * hn * hn
    ab = np.zeros((3, n))
    ab[0, 1:] = up[:n - 1]
    ab[1, :] = diag
    ab[2, :n - 1] = lo[1:]
    return solve_banded((1, 1), ab, rhs)


def _sort_key_order(lm_l: np.ndarray, lm_m: np.ndarray) -> np.ndarray:
    # primary |m|, then l, then cos (m>=0) before sin (m<0)
    return np.lexsort((lm_m < 0, lm_l, np.abs(lm_m)))


@dataclass
class LegendreProgram:
    """Per-potential recurrence program consumed by the Multipole kernels.

    Every step is P_l = alpha (c + beta) P_{l-s} + gamma P_{l-2s}.  A group
    whose active terms all have l - |m| even steps l by s = 2 with c = cos^2
    (the same-parity recurrence); the others step by s = 1 with c = cos, beta
    = 0 and gamma = minus the one-step recurrence's beta.  (alpha times a sum
    rather than alpha c + beta: one coefficient per instruction, the most a
    gfx950 FP64 instruction takes from scalar registers.)"""
    grp_m: np.ndarray        # int32 (G,)
    grp_step0: np.ndarray    # int32 (G+1,)  steps of group g: [step0[g], step0[g+1])
    qmm: np.ndarray          # float64 (G,)  mul_m * normalised P_m^m / sin^m
    step_alpha: np.ndarray   # float64 (S,)
    step_beta: np.ndarray    # float64 (S,)
    step_cos: np.ndarray     # int32 (S,)  table column of the cos term, -1 if none
    step_sin: np.ndarray     # int32 (S,)
    step_l: np.ndarray       # int32 (S,)
    grp_stride: np.ndarray   # int32 (G,): 1 or 2
    step_gamma: np.ndarray   # float64 (S,)


def _qmm(m: int) -> float:
    """mul_m * sqrt((2m+1)/(4 pi (2m)!)) (-1)^m (2m-1)!!, by the stable product
    P_mm = -sqrt((2m+1)/(2m)) sin(theta) P_{m-1,m-1}, P_00 = 1/sqrt(4 pi)."""
    v = 1.0 / math.sqrt(4.0 * math.pi)
    for j in range(1, m + 1):
        v *= -math.sqrt((2.0 * j + 1.0) / (2.0 * j))
    mul = 2.0 * math.sqrt(math.pi) if m == 0 else 2.0 * math.sqrt(2.0 * math.pi)
    return mul * v


def _recurrence_ab(l: int, am: int):
    """alpha_l, beta_l of the normalised one-step recurrence
    P_l = alpha_l cos P_{l-1} - beta_l P_{l-2} (l > |m|)."""
    if l == am + 1:
        return math.sqrt(2.0 * am + 3.0), 0.0
    l2, m2 = float(l * l), float(am * am)
    return (math.sqrt((4.0 * l2 - 1.0) / (l2 - m2)),
            math.sqrt((2.0 * l + 1.0) * ((l - 1.0) ** 2 - m2) / ((2.0 * l - 3.0) * (l2 - m2))))


def _recurrence2(l: int, am: int):
    """(alpha, beta, gamma) of the two-step recurrence
    P_l = (alpha cos^2 + beta) P_{l-2} + gamma P_{l-4}, l >= |m| + 2: the
    one-step recurrence applied twice, with cos P_{l-3} eliminated through
    P_{l-2} = alpha_{l-2} cos P_{l-3} - beta_{l-2} P_{l-4}."""
    a_l, b_l = _recurrence_ab(l, am)
    a_1, b_1 = _recurrence_ab(l - 1, am)
    if l == am + 2:
        return a_l * a_1, -b_l, 0.0
    a_2, b_2 = _recurrence_ab(l - 2, am)
    return a_l * a_1, -(a_l * b_1 / a_2 + b_l), -a_l * b_1 * b_2 / a_2


def build_legendre_program(lm_l: np.ndarray, lm_m: np.ndarray) -> LegendreProgram:
    """Group the sorted active harmonics (column 0 = (0,0) excluded: the
    kernels treat the monopole separately) into |m| runs; a run whose terms
    all have l - |m| even steps by 2 (the skipped odd steps carry no term)."""
    cols: Dict[int, Dict[int, List[int]]] = {}
    for c in range(1, len(lm_l)):
        l, m = int(lm_l[c]), int(lm_m[c])
        am = abs(m)
        slot = cols.setdefault(am, {}).setdefault(l, [-1, -1])
        slot[0 if m >= 0 else 1] = c
    grp_m, grp_step0, qmm, stride = [], [0], [], []
    alpha, beta, gamma, scos, ssin, sl = [], [], [], [], [], []
    for am in sorted(cols):
        ls = cols[am]
        l_last = max(ls)
        two = all((l - am) % 2 == 0 for l in ls)
        grp_m.append(am)
        qmm.append(_qmm(am))
        stride.append(2 if two else 1)
        for l in range(am, l_last + 1, 2 if two else 1):
            if l == am:
                a, b, g = 0.0, 0.0, 0.0
            elif two:
                a, b, g = _recurrence2(l, am)
                b /= a
            else:
                a, b1 = _recurrence_ab(l, am)
                b, g = 0.0, -b1
            c_cos, c_sin = ls.get(l, [-1, -1])
            alpha.append(a)
            beta.append(b)
            gamma.append(g)
            scos.append(c_cos)
            ssin.append(c_sin)
            sl.append(l)
        grp_step0.append(len(alpha))
    return LegendreProgram(
        np.asarray(grp_m, np.int32), np.asarray(grp_step0, np.int32),
        np.asarray(qmm, np.float64), np.asarray(alpha, np.float64),
        np.asarray(beta, np.float64), np.asarray(scos, np.int32),
        np.asarray(ssin, np.int32), np.asarray(sl, np.int32),
        np.asarray(stride, np.int32), np.asarray(gamma, np.float64))


def triaxial_degree(prog: LegendreProgram) -> int:
    """L when the program is the triaxial program of even degree L >= 2:
    groups m = 0, 2, ..., L, group m stepping l = m, m + 2, ..., L by 2, cos
    terms only, the cos term of step j in column j for j >= 1 (step 0, l = m =
    0, has none: the monopole is evaluated apart) -- the program of a table
    with every harmonic that triaxial symmetry allows up to degree L.
    0 otherwise."""
    G = len(prog.grp_m)
    L = 2 * (G - 1)
    if G < 2 or not np.array_equal(prog.grp_m, np.arange(0, L + 1, 2)):
        return 0
    step_l = np.concatenate([np.arange(m, L + 1, 2) for m in range(0, L + 1, 2)])
    S = len(step_l)
    step0 = np.concatenate([[0], np.cumsum([(L - m) // 2 + 1 for m in range(0, L + 1, 2)])])
    ok = (np.array_equal(prog.grp_step0, step0) and np.array_equal(prog.step_l, step_l)
          and np.array_equal(prog.step_cos, np.concatenate([[-1], np.arange(1, S)]))
          and np.all(prog.step_sin < 0) and np.all(prog.grp_stride == 2))
    return L if ok else 0


def full_degree(prog: LegendreProgram) -> int:
    """L when the program is the full program of degree L >= 1: groups
    m = 0, 1, .., L, group m stepping l = m, m + 1, .., L by 1, the cos term
    of step j in column j for 1 <= j <= L (step 0, l = m = 0, has none: the
    monopole is evaluated apart) and for a later step j the cos term in
    column 2j - L - 1 and the sin term in column 2j - L -- the program of a
    table with every harmonic up to degree L.  0 otherwise."""
    G = len(prog.grp_m)
    L = G - 1
    if G < 2 or not np.array_equal(prog.grp_m, np.arange(L + 1)):
        return 0
    step_l = np.concatenate([np.arange(m, L + 1) for m in range(L + 1)])
    S = len(step_l)
    step0 = np.concatenate([[0], np.cumsum([L - m + 1 for m in range(L + 1)])])
    j = np.arange(S)
    cos = np.where(j <= L, j, 2 * j - L - 1)
    cos[0] = -1
    sin = np.where(j <= L, -1, 2 * j - L)
    ok = (np.array_equal(prog.grp_step0, step0) and np.array_equal(prog.step_l, step_l)
          and np.array_equal(prog.step_cos, cos) and np.array_equal(prog.step_sin, sin)
          and np.all(prog.grp_stride[np.diff(prog.grp_step0) > 1] == 1))
    return L if ok else 0


def program_flags(prog: LegendreProgram) -> int:
    """agx_multipole.prog_flags for this program: AGX_PROG_STRIDE1 (1) /
    AGX_PROG_STRIDE2 (4) when every group that recurs (more than one step)
    has stride 1 / 2, plus AGX_PROG_COS_ONLY (2) when no step has a sin term,
    plus the triaxial degree (triaxial_degree) in bits 8-15 and the full
    degree (full_degree) in bits 16-23."""
    steps = np.diff(prog.grp_step0)
    recurs = steps > 1
    stride1 = not np.any((prog.grp_stride != 1) & recurs)
    stride2 = not np.any((prog.grp_stride != 2) & recurs)
    cos_only = not np.any(prog.step_sin >= 0)
    return (int(stride1) | (2 if cos_only else 0) | (4 if stride2 else 0)
            | (triaxial_degree(prog) << 8) | (full_degree(prog) << 16))


# 32 bytes: four int32 (group records: m, step_begin, step_end, stride; step
# records: l, the packed offsets and, over i2-i3, the double gamma), then two
# doubles (group: qmm, m; step: alpha, beta).  The step record of l = m is all
# coefficients zero.
PROGRAM_RECORD = np.dtype({"names": ["i0", "i1", "i2", "i3", "g", "d0", "d1"],
                           "formats": ["<i4", "<i4", "<i4", "<i4", "<f8", "<f8", "<f8"],
                           "offsets": [0, 4, 8, 12, 8, 16, 24], "itemsize": 32})


def pack_offsets(cos_off: np.ndarray, sin_off: np.ndarray) -> np.ndarray:
    """int32 holding the cos offset (int16, low half) and the sin offset (int16,
    high half); -1 marks an absent term."""
    lo = cos_off.astype(np.int64) & 0xFFFF
    hi = (sin_off.astype(np.int64) & 0xFFFF) << 16
    return (lo | hi).astype(np.uint32).view(np.int32)


def unpack_offsets(packed: np.ndarray):
    p = np.asarray(packed, np.int32)
    return (p.astype(np.uint32) & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.int32), \
        (p >> 16).astype(np.int32)


def pack_program(prog: LegendreProgram) -> np.ndarray:
    """The kernels' 32-byte program records (layout: agx_multipole.program in
    include/agama_gpu.h): group records, a zero record, step records (with
    column offsets 6*c), a zero record."""
    G, S = len(prog.grp_m), len(prog.step_l)
    if S and 6 * (max(int(prog.step_cos.max()), int(prog.step_sin.max())) + 1) > 32767:
        raise ValueError("too many terms for 16-bit slab offsets")
    rec = np.zeros(G + S + 2, PROGRAM_RECORD)
    g = rec[:G]
    g["i0"] = prog.grp_m
    g["i1"] = prog.grp_step0[:-1]
    g["i2"] = prog.grp_step0[1:]
    g["i3"] = prog.grp_stride
    g["d0"] = prog.qmm
    g["d1"] = prog.grp_m
    st = rec[G + 1:G + 1 + S]
    st["i0"] = prog.step_l
    st["i1"] = pack_offsets(np.where(prog.step_cos >= 0, 6 * prog.step_cos, -1),
                            np.where(prog.step_sin >= 0, 6 * prog.step_sin, -1))
    st["g"] = prog.step_gamma
    st["d0"] = prog.step_alpha
    st["d1"] = prog.step_beta
    return rec


def build_multipole_data(coefs, prune_threshold: float = 1e-16) -> dict:
    """MultipoleCoefs -> reference-layout tables plus the kernel layout.

    Keys follow the reference (gpu_potential.py:607-626): poly (n_lm, n_int, 6),
    logr_min, dlogr, inv_dlogr, n_intervals, n_lm, lmax, lm_l, lm_m, r_grid,
    log_scaling, invPhi0, inner_{s,U,W}, outer_{s,U,W}; plus the log-mode
    flags inner_log / outer_log (see ``inner_extrapolation``), ``poly_km``
    (n_int, n_lm, 6) and ``program`` (LegendreProgram).
    """
    if coefs.dphi_dr is None:
        raise ValueError(
            "MultipoleCoefs.dphi_dr is None; dPhi/dr data is required.\n"
            "Use a .coef_mul_DR file (not plain .coef_mul).")
    r = np.asarray(coefs.R_grid, np.float64).copy()
    phi = np.asarray(coefs.phi, np.float64).copy()
    dphi_dr = np.asarray(coefs.dphi_dr, np.float64).copy()
    nR, n_lm = phi.shape
    if nR < 2:
        raise ValueError(f"Need at least 2 radial grid points, got {nR}.")
    if len(coefs.lm_labels) != n_lm:
        raise ValueError("lm_labels does not match the number of columns")

    logr = np.log(r)
    steps = np.diff(logr)
    dlogr = float(steps.mean())
    spread = float(np.max(np.abs(steps - dlogr)) / dlogr)
    dphi_dlogr = dphi_dr * r[:, None]
    d0_dr = dphi_dr[:, 0]
    if spread > 1e-3:
        # resample onto a log-uniform grid of >= 1000 nodes with a cubic
        # Hermite interpolant (gpu_potential.py:468-489)
        from scipy.interpolate import CubicHermiteSpline
        n_new = max(nR, 1000)
        logr_new = np.linspace(logr[0], logr[-1], n_new)
        r_new = np.exp(logr_new)
        spl = CubicHermiteSpline(logr, phi, dphi_dlogr)
        phi = spl(logr_new)
        dphi_dlogr = spl(logr_new, 1)
        d0_dr = dphi_dlogr[:, 0] / r_new
        r, logr, nR = r_new, logr_new, n_new
        steps = np.diff(logr)
        dlogr = float(steps.mean())
    elif spread > 1e-6:
        warnings.warn(
            f"Radial grid is not perfectly log-uniform (max relative spacing "
            f"error = {spread:.2e}). O(1) grid lookup may be slightly inaccurate.",
            stacklevel=3)

    logr_min = float(logr[0])
    inv_dlogr = 1.0 / dlogr
    n_int = nR - 1
    outer_s, outer_U, outer_W, outer_log = outer_extrapolation(phi[:, 0], d0_dr, r)

    lm_l = np.array([l for l, _ in coefs.lm_labels], np.int32)
    lm_m = np.array([m for _, m in coefs.lm_labels], np.int32)
    lmax = int(lm_l.max())
    if lmax > LMAX_LIMIT:
        raise ValueError(f"lmax={lmax} exceeds the kernel's compile-time limit of {LMAX_LIMIT}.")
    order = _sort_key_order(lm_l, lm_m)
    columns = order.copy()
    lm_l, lm_m = lm_l[order], lm_m[order]
    phi, dphi_dlogr = phi[:, order], dphi_dlogr[:, order]
    if not (lm_l[0] == 0 and lm_m[0] == 0):
        raise ValueError("the (l=0, m=0) harmonic is required")

    if prune_threshold > 0:
        keep = np.ones(n_lm, bool)
        keep[1:] = np.any(np.abs(phi[:, 1:]) >= prune_threshold, axis=0)
        if not keep.all():
            phi, dphi_dlogr = phi[:, keep], dphi_dlogr[:, keep]
            columns = columns[keep]
            lm_l, lm_m = lm_l[keep], lm_m[keep]
            n_lm = int(keep.sum())
            lmax = int(lm_l.max())

    p0 = phi[:, 0]
    log_scaling = bool(np.all(p0 < 0.0))
    invPhi0, inner_s, inner_U, inner_W, inner_log = 0.0, 0.0, 0.0, float(p0[0]), False
    if log_scaling:
        invPhi0, inner_s, inner_U, inner_W, inner_log = inner_extrapolation(p0, d0_dr, r)

    f = phi.copy()
    df = dphi_dlogr.copy()
    if log_scaling:
        f[:, 0] = np.log(invPhi0 - 1.0 / p0)
        df[:, 0] = dphi_dlogr[:, 0] / (p0 * (invPhi0 * p0 - 1.0))
        ratio = phi[:, 1:] / p0[:, None]
        f[:, 1:] = ratio
        df[:, 1:] = (dphi_dlogr[:, 1:] - ratio * dphi_dlogr[:, 0:1]) / p0[:, None]
    d2 = quintic_second_derivs(logr, f, df)

    h = dlogr
    f0, f1 = f[:-1], f[1:]
    m0, m1 = df[:-1] * h, df[1:] * h
    q0, q1 = d2[:-1] * h * h, d2[1:] * h * h
    dlt = f1 - f0
    coef = np.stack([
        f0,
        m0,
        0.5 * q0,
        10.0 * dlt - 6.0 * m0 - 4.0 * m1 - 1.5 * q0 + 0.5 * q1,
        -15.0 * dlt + 8.0 * m0 + 7.0 * m1 + 1.5 * q0 - q1,
        6.0 * dlt - 3.0 * m0 - 3.0 * m1 + 0.5 * (q1 - q0),
    ], axis=-1)                                   # (n_int, n_lm, 6)
    poly_km = np.ascontiguousarray(coef, dtype=np.float64)
    poly = np.ascontiguousarray(coef.transpose(1, 0, 2), dtype=np.float64)

    return dict(
        poly=poly, poly_km=poly_km,
        logr_min=logr_min, dlogr=dlogr, inv_dlogr=inv_dlogr,
        n_intervals=n_int, n_lm=n_lm, lmax=lmax,
        lm_l=lm_l.astype(np.int32), lm_m=lm_m.astype(np.int32), r_grid=r,
        log_scaling=int(log_scaling), invPhi0=float(invPhi0),
        inner_s=float(inner_s), inner_U=float(inner_U), inner_W=float(inner_W),
        inner_log=int(inner_log),
        outer_s=float(outer_s), outer_U=float(outer_U), outer_W=float(outer_W),
        outer_log=int(outer_log),
        program=build_legendre_program(lm_l, lm_m),
        # structure for a differentiable rebuild of the same tables
        columns=columns.astype(np.int64), resampled=bool(spread > 1e-3),
        inner_mode=("log" if inner_log else ("s2" if inner_s == 2.0 else "const")),
        inner_reset=bool(inner_s == 2.0 and invPhi0 == 0.0),
        outer_mode=("log" if outer_log else "kepler"),
    )
