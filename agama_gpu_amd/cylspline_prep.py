"""Host-side CylSpline table builder (CPU, once per potential).

Restates the reference's ``_build_cylspline_data`` (gpu_potential.py:
1143-1253) and helpers ``_sph_harm_agama`` (963-998),
``_determine_asympt_cylspline`` (1001-1095), ``_setup_cubic2d_nodes``
(1098-1140): z-mirroring of half-space grids, the outer PowerLaw multipole
fit on the grid boundary (lmax 8), Rscale = -Mtot/Phi0, asinh-scaled grids,
log-scaling of the m=0 harmonic, and bicubic Hermite node data
(f, df/dlR, df/dlz, d2f/dlR dlz) per harmonic.
"""
from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy as np

LMAX_OUTER = 8
MUL0 = 2.0 * math.sqrt(math.pi)
MUL1 = 2.0 * math.sqrt(2.0 * math.pi)


def ptilde(lmax: int, m: int, ct: float, st: float) -> np.ndarray:
    """Agama-normalised P_l^m(cos theta) for l = m..lmax (Condon-Shortley),
    Ptilde = sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) P_l^m."""
    out = np.zeros(lmax - m + 1)
    v = 1.0 / math.sqrt(4.0 * math.pi)
    for j in range(1, m + 1):
        v *= -math.sqrt((2.0 * j + 1.0) / (2.0 * j)) * st
    out[0] = v
    if lmax == m:
        return out
    out[1] = math.sqrt(2.0 * m + 3.0) * ct * v
    for l in range(m + 2, lmax + 1):
        l2, m2 = float(l * l), float(m * m)
        a = math.sqrt((4.0 * l2 - 1.0) / (l2 - m2))
        b = math.sqrt((2.0 * l + 1.0) * ((l - 1.0) ** 2 - m2) / ((2.0 * l - 3.0) * (l2 - m2)))
        out[l - m] = a * ct * out[l - m - 1] - b * out[l - m - 2]
    return out


def outer_asymptote(R: np.ndarray, z: np.ndarray, phi: Dict[int, np.ndarray], mmax: int,
                    lmax_fit: int = LMAX_OUTER) -> Tuple[np.ndarray, float]:
    """Least-squares fit of Phi = sum W_lm (r/r0)^-(l+1) mul Ptilde_lm T_m to the
    grid's outer boundary nodes (gpu_potential.py:1001-1095).
    Returns W[(lmax_fit+1)^2] indexed l(l+1)+m and r0."""
    from scipy.linalg import lstsq
    nR, nz = len(R), len(z)
    zsym = z[0] == 0
    pts = []
    for iR in range(nR - 1):
        pts.append((R[iR], z[nz - 1], iR, nz - 1))
        pts.append((R[iR], -z[nz - 1], iR, nz - 1) if zsym else (R[iR], z[0], iR, 0))
    for iz in range(nz):
        pts.append((R[nR - 1], z[iz], nR - 1, iz))
        if zsym and iz > 0:
            pts.append((R[nR - 1], -z[iz], nR - 1, iz))
    Rp = np.array([p[0] for p in pts])
    zp = np.array([p[1] for p in pts])
    iRp = np.array([p[2] for p in pts])
    izp = np.array([p[3] for p in pts])
    rp = np.hypot(Rp, zp)
    ct, st = zp / rp, Rp / rp
    r0 = min(float(R[-1]), float(np.max(np.abs(z))))
    W = np.zeros((lmax_fit + 1) ** 2)
    mfit = min(lmax_fit, mmax)
    for m in range(-mfit, mfit + 1):
        if m not in phi:
            continue
        am = abs(m)
        ncol = lmax_fit - am + 1
        mul = MUL0 if m == 0 else MUL1
        M = np.zeros((len(pts), ncol))
        rhs = phi[m][iRp, izp]
        for p in range(len(pts)):
            P = ptilde(lmax_fit, am, ct[p], st[p])
            for li in range(ncol):
                M[p, li] = P[li] * (rp[p] / r0) ** (-(am + li + 1)) * mul
        sol = lstsq(M, rhs)[0]
        for li in range(ncol):
            l = am + li
            W[l * (l + 1) + m] = sol[li]
    if not np.isfinite(W[0]):
        p0 = phi.get(0)
        avg = float(np.mean(p0[iRp, izp] * rp / r0)) if p0 is not None else 0.0
        W[:] = 0.0
        W[0] = avg
    return W, r0


def cubic2d_nodes(lR: np.ndarray, lz: np.ndarray, f: np.ndarray) -> np.ndarray:
    """(f, f_x, f_y, f_xy) per node of a 2-D cubic spline: natural in lz;
    in lR zero slope at the first node and natural at the last
    (gpu_potential.py:1098-1140).  Returns (nR, nz, 4)."""
    from scipy.interpolate import CubicSpline
    nR, nz = f.shape
    fy = np.empty_like(f)
    fx = np.empty_like(f)
    fxy = np.empty_like(f)
    for i in range(nR):
        fy[i] = CubicSpline(lz, f[i], bc_type="natural")(lz, 1)
    for j in range(nz):
        fx[:, j] = CubicSpline(lR, f[:, j], bc_type=((1, 0.0), "natural"))(lR, 1)
        fxy[:, j] = CubicSpline(lR, fy[:, j], bc_type=((1, 0.0), "natural"))(lR, 1)
    return np.stack([f, fx, fy, fxy], axis=-1)


def build_cylspline_data(coefs) -> dict:
    """CylSplineCoefs -> device tables and scalars (gpu_potential.py:1143-1253)."""
    R = np.asarray(coefs.R_grid, np.float64)
    z0 = np.asarray(coefs.z_grid, np.float64)
    src = {int(m): np.asarray(v, np.float64) for m, v in coefs.phi.items()}
    if 0 not in src:
        raise ValueError("CylSpline coefficients need the m=0 harmonic")
    nR = len(R)
    if nR < 2 or len(z0) < 2:
        raise ValueError("CylSpline grid needs at least 2 nodes in R and z")
    mmax = max(abs(m) for m in src)
    if z0[0] == 0:          # half-space grid: mirror to z < 0
        z = np.concatenate([-z0[:0:-1], z0])
        n0 = len(z0)
        tabs = {}
        for m, t in src.items():
            full = np.empty((nR, 2 * n0 - 1))
            full[:, n0 - 1:] = t
            full[:, :n0] = t[:, ::-1]
            tabs[m] = full
        Phi0 = tabs[0][0, n0 - 1]
    else:
        z = z0
        tabs = src
        # reference: index (nz+1)//2 of the m=0 table at R=0 (gpu_potential.py:1188-1190)
        Phi0 = tabs[0][0, (len(z) + 1) // 2]
    W, r0 = outer_asymptote(R, z, tabs, mmax)
    Mtot = -W[0] * r0 * MUL0
    Rscale = -Mtot / Phi0 if (Phi0 < 0.0 and Mtot > 0.0) else float(R[nR // 2])
    lR = np.arcsinh(R / Rscale)
    lz = np.arcsinh(z / Rscale)
    log_scaling = bool(np.all(tabs[0] < 0.0))
    # harmonics ordered by (|m|, cos before sin): the kernels advance
    # cos/sin(|m| phi) by one rotation per new |m|
    ms = sorted(tabs, key=lambda m: (abs(m), m < 0))
    nodes = []
    for m in ms:
        f = tabs[m]
        if log_scaling:
            f = np.log(-f) if m == 0 else f / tabs[0]
        nodes.append(cubic2d_nodes(lR, lz, f))
    node_arr = np.ascontiguousarray(np.stack(nodes, axis=0), dtype=np.float64)
    return dict(Rscale=float(Rscale), lR_grid=lR, lz_grid=lz, node_arr=node_arr,
                m_arr=np.asarray(ms, np.int32), n_harm=len(ms), mmax=int(mmax),
                log_scaling=int(log_scaling), W_outer=W, r0_outer=float(r0),
                lmax_outer=LMAX_OUTER, mmax_outer=min(LMAX_OUTER, mmax))
