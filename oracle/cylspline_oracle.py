"""ORACLE -- test infrastructure only (numpy restatement of the reference's
CylSpline path).

Table build (gpu_potential.py:963-1253): z mirroring, outer PowerLaw fit on
the boundary nodes, Rscale, asinh grids, log-scaling, 2-D cubic-spline node
derivatives.  The cubic splines are this file's own tridiagonal solves (the
product uses scipy.interpolate.CubicSpline) and the fit uses numpy's lstsq.
The associated Legendre functions come from scipy.special.sph_harm_y.

Evaluation (_cylspl_potential_kernel.cu:396-863): per-harmonic bicubic
Hermite interpolation in (lR, lz) with the corner offset, direct cos/sin(m
phi), log-scaling unscale, cylindrical -> Cartesian chain rule.  Corrections
relative to the reference, as in the product: the Hessian keeps the
dPhi/dphi terms, and outside the grid the PowerLaw expansion is evaluated in
full -- here its gradient comes from complex-step differentiation and its
Hessian from central differences of that gradient, independently of any
hand-derived chain rule.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import sph_harm_y

G = 4.30091727067736e-06
INV_4PIG = 1.0 / (4.0 * math.pi * G)
MUL0, MUL1 = 2.0 * math.sqrt(math.pi), 2.0 * math.sqrt(2.0 * math.pi)
LMAX_O = 8


def spline_slopes(x, f, left_slope=None):
    """First derivatives at the nodes of the C2 cubic spline through (x, f):
    natural at both ends, or clamped slope ``left_slope`` at x[0] and natural
    at x[-1].  Tridiagonal system in the node slopes (Thomas algorithm)."""
    n = len(x)
    h = np.diff(x)
    A = np.zeros((n, 3))          # sub, diag, sup
    b = np.zeros(n)
    d = np.diff(f) / h
    for i in range(1, n - 1):
        A[i] = (1.0 / h[i - 1], 2.0 / h[i - 1] + 2.0 / h[i], 1.0 / h[i])
        b[i] = 3.0 * (d[i - 1] / h[i - 1] + d[i] / h[i])
    if left_slope is None:
        A[0] = (0.0, 2.0 / h[0], 1.0 / h[0])
        b[0] = 3.0 * d[0] / h[0]
    else:
        A[0] = (0.0, 1.0, 0.0)
        b[0] = left_slope
    A[n - 1] = (1.0 / h[-1], 2.0 / h[-1], 0.0)
    b[n - 1] = 3.0 * d[-1] / h[-1]
    c = np.zeros(n)
    bb = b.copy()
    dd = A[:, 1].copy()
    for i in range(1, n):
        w = A[i, 0] / dd[i - 1]
        dd[i] -= w * A[i - 1, 2]
        bb[i] -= w * bb[i - 1]
    c[n - 1] = bb[n - 1] / dd[n - 1]
    for i in range(n - 2, -1, -1):
        c[i] = (bb[i] - A[i, 2] * c[i + 1]) / dd[i]
    return c


def _ptilde_scipy(l, m, ct):
    th = np.arccos(np.clip(ct, -1, 1))
    return np.real(sph_harm_y(l, m, th, 0.0 * th))


class OracleCylSpline:
    def __init__(self, coefs):
        R = np.asarray(coefs.R_grid, float)
        z0 = np.asarray(coefs.z_grid, float)
        src = {int(m): np.asarray(v, float) for m, v in coefs.phi.items()}
        nR = len(R)
        mmax = max(abs(m) for m in src)
        if z0[0] == 0:
            z = np.concatenate([-z0[:0:-1], z0])
            n0 = len(z0)
            tabs = {m: np.concatenate([t[:, ::-1], t[:, 1:]], axis=1) for m, t in src.items()}
            Phi0 = tabs[0][0, n0 - 1]
        else:
            z, tabs = z0, src
            Phi0 = tabs[0][0, (len(z) + 1) // 2]
        self.W, self.r0 = self._fit(R, z, tabs, mmax)
        Mtot = -self.W[0] * self.r0 * MUL0
        self.Rscale = -Mtot / Phi0 if (Phi0 < 0 and Mtot > 0) else float(R[nR // 2])
        self.lR = np.arcsinh(R / self.Rscale)
        self.lz = np.arcsinh(z / self.Rscale)
        self.log_scaling = bool(np.all(tabs[0] < 0))
        self.mmax_o = min(LMAX_O, mmax)
        self.harm = {}
        for m, t in tabs.items():
            f = (np.log(-t) if m == 0 else t / tabs[0]) if self.log_scaling else t
            fy = np.stack([spline_slopes(self.lz, f[i]) for i in range(nR)])
            fx = np.stack([spline_slopes(self.lR, f[:, j], 0.0) for j in range(len(z))], axis=1)
            fxy = np.stack([spline_slopes(self.lR, fy[:, j], 0.0) for j in range(len(z))], axis=1)
            self.harm[m] = (f, fx, fy, fxy)

    @staticmethod
    def _fit(R, z, tabs, mmax):
        nR, nz = len(R), len(z)
        zsym = z[0] == 0
        pts = [(R[i], z[nz - 1], i, nz - 1) for i in range(nR - 1)]
        pts += [(R[i], -z[nz - 1], i, nz - 1) if zsym else (R[i], z[0], i, 0) for i in range(nR - 1)]
        pts += [(R[nR - 1], z[j], nR - 1, j) for j in range(nz)]
        if zsym:
            pts += [(R[nR - 1], -z[j], nR - 1, j) for j in range(1, nz)]
        P = np.array(pts)
        Rp, zp, iR, iz = P[:, 0], P[:, 1], P[:, 2].astype(int), P[:, 3].astype(int)
        rp = np.hypot(Rp, zp)
        ct = zp / rp
        r0 = min(float(R[-1]), float(np.max(np.abs(z))))
        W = np.zeros((LMAX_O + 1) ** 2)
        mf = min(LMAX_O, mmax)
        for m in range(-mf, mf + 1):
            if m not in tabs:
                continue
            am = abs(m)
            mul = MUL0 if m == 0 else MUL1
            cols = [_ptilde_scipy(l, am, ct) * (rp / r0) ** (-(l + 1)) * mul for l in range(am, LMAX_O + 1)]
            sol = np.linalg.lstsq(np.stack(cols, axis=1), tabs[m][iR, iz], rcond=None)[0]
            for k, l in enumerate(range(am, LMAX_O + 1)):
                W[l * (l + 1) + m] = sol[k]
        return W, r0

    # ---- outer PowerLaw multipole (complex-step safe) -----------------------
    def _outer_phi(self, x, y, z):
        R = np.sqrt(x * x + y * y)
        r = np.sqrt(R * R + z * z)
        ct, st = z / r, R / r
        safe = np.abs(R) > 0
        cph = np.where(safe, x / np.where(safe, R, 1.0), 1.0)
        sph = np.where(safe, y / np.where(safe, R, 1.0), 0.0)
        q = self.r0 / r
        total = 0.0 * x
        cm, sm = 1.0 + 0.0 * x, 0.0 * x
        pmm = 1.0 / math.sqrt(4 * math.pi) + 0.0 * x
        for m in range(0, self.mmax_o + 1):
            if m > 0:
                cm, sm = cm * cph - sm * sph, sm * cph + cm * sph
                pmm = -math.sqrt((2 * m + 1) / (2 * m)) * st * pmm
            mul = MUL0 if m == 0 else MUL1
            p2, p1 = 0.0 * x, pmm
            for l in range(m, LMAX_O + 1):
                if l == m + 1:
                    p2, p1 = p1, math.sqrt(2 * m + 3) * ct * p1
                elif l > m + 1:
                    a = math.sqrt((4 * l * l - 1) / (l * l - m * m))
                    b = math.sqrt((2 * l + 1) * ((l - 1) ** 2 - m * m) / ((2 * l - 3) * (l * l - m * m)))
                    p2, p1 = p1, a * ct * p1 - b * p2
                rad = q ** (l + 1) * mul * p1
                total = total + self.W[l * (l + 1) + m] * rad * cm
                if m > 0:
                    total = total + self.W[l * (l + 1) - m] * rad * sm
        return total

    def _outer(self, x):
        X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
        P = self._outer_phi(X, Y, Z)
        hcs = 1e-30

        def grad(xx):
            out = []
            for k in range(3):
                c = [xx[:, 0].astype(complex), xx[:, 1].astype(complex), xx[:, 2].astype(complex)]
                c[k] = c[k] + 1j * hcs
                out.append(np.imag(self._outer_phi(*c)) / hcs)
            return np.stack(out, axis=1)
        g = grad(x)
        step = 1e-4 * np.linalg.norm(x, axis=1)
        H = np.zeros((len(x), 6))
        gs = []
        for k in range(3):
            e = np.zeros(3)
            e[k] = 1.0
            gs.append((grad(x + step[:, None] * e) - grad(x - step[:, None] * e)) / (2 * step[:, None]))
        for s, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
            H[:, s] = 0.5 * (gs[j][:, i] + gs[i][:, j])
        return P, g, H

    # ---- evaluation -----------------------------------------------------------
    def eval(self, xyz):
        """(phi, +grad Phi, Hessian [xx,yy,zz,xy,yz,xz], outer mask)."""
        x = np.asarray(xyz, float).reshape(-1, 3)
        X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
        R = np.hypot(X, Y)
        lR = np.arcsinh(R / self.Rscale)
        lz = np.arcsinh(Z / self.Rscale)
        outer = (lR < self.lR[0]) | (lR > self.lR[-1]) | (lz < self.lz[0]) | (lz > self.lz[-1])
        n = len(x)
        phi = np.zeros(n)
        g = np.zeros((n, 3))
        H = np.zeros((n, 6))
        if np.any(outer):
            phi[outer], g[outer], H[outer] = self._outer(x[outer])
        ins = ~outer
        if np.any(ins):
            p_, g_, h_ = self._inner(x[ins], R[ins], lR[ins], lz[ins])
            phi[ins], g[ins], H[ins] = p_, g_, h_
        return phi, g, H, outer

    def _inner(self, x, R, lR, lz):
        nR, nz = len(self.lR), len(self.lz)
        iR = np.clip(np.searchsorted(self.lR, lR, side="right") - 1, 0, nR - 2)
        iz = np.clip(np.searchsorted(self.lz, lz, side="right") - 1, 0, nz - 2)
        xl, xh = self.lR[iR], self.lR[iR + 1]
        yl, yh = self.lz[iz], self.lz[iz + 1]
        phi_ang = np.arctan2(x[:, 1], x[:, 0])

        def c1(t, h, fl, fh, dl, dh):        # evalCubic1D, :110-128
            T = 1 - t
            tT = t * T
            dif = fh - fl
            Q = 3 * (dl + dh) - 6 * dif / h
            return (fl * T + fh * t + (dif * (t - T) + (dl * T - dh * t) * h) * tT,
                    dl * T + dh * t - Q * tT, (dh - dl + Q * (t - T)) / h)

        tz, hz = (lz - yl) / (yh - yl), yh - yl
        tR, hR = (lR - xl) / (xh - xl), xh - xl
        S = np.ones(len(x)) if self.log_scaling else np.zeros(len(x))
        acc = {k: np.zeros(len(x)) for k in ("R", "Z", "RR", "RZ", "ZZ", "p", "pp", "Rp", "Zp")}
        F0 = None
        q2 = {}
        for m, (f, fx, fy, fxy) in self.harm.items():
            fll, flu, ful, fuu = f[iR, iz], f[iR, iz + 1], f[iR + 1, iz], f[iR + 1, iz + 1]
            off = np.where(lR == xh, np.where(lz == yh, fuu, ful), np.where(lz == yh, flu, fll))
            flow = [fll - off, ful - off, fx[iR, iz], fx[iR + 1, iz]]
            fupp = [flu - off, fuu - off, fx[iR, iz + 1], fx[iR + 1, iz + 1]]
            dflow = [fy[iR, iz], fy[iR + 1, iz], fxy[iR, iz], fxy[iR + 1, iz]]
            dfupp = [fy[iR, iz + 1], fy[iR + 1, iz + 1], fxy[iR, iz + 1], fxy[iR + 1, iz + 1]]
            Fz = [c1(tz, hz, flow[k], fupp[k], dflow[k], dfupp[k]) for k in range(4)]
            v, vR, vRR = c1(tR, hR, Fz[0][0], Fz[1][0], Fz[2][0], Fz[3][0])
            v = v + off
            vZ, vRZ, _ = c1(tR, hR, Fz[0][1], Fz[1][1], Fz[2][1], Fz[3][1])
            vZZ, _, _ = c1(tR, hR, Fz[0][2], Fz[1][2], Fz[2][2], Fz[3][2])
            if m in (2, -2):
                q2[m] = (v, vR, vRR)
            if m == 0:
                F0 = (v, vR, vZ, vRR, vRZ, vZZ)
                if not self.log_scaling:
                    S = S + v
                    for k, val in zip(("R", "Z", "RR", "RZ", "ZZ"), (vR, vZ, vRR, vRZ, vZZ)):
                        acc[k] = acc[k] + val
                continue
            am = abs(m)
            T = np.cos(am * phi_ang) if m > 0 else np.sin(am * phi_ang)
            dT = -am * np.sin(am * phi_ang) if m > 0 else am * np.cos(am * phi_ang)
            S = S + v * T
            acc["R"] += vR * T
            acc["Z"] += vZ * T
            acc["RR"] += vRR * T
            acc["RZ"] += vRZ * T
            acc["ZZ"] += vZZ * T
            acc["p"] += v * dT
            acc["pp"] += -am * am * v * T
            acc["Rp"] += vR * dT
            acc["Zp"] += vZ * dT
        v0, v0R, v0Z, v0RR, v0RZ, v0ZZ = F0
        if self.log_scaling:
            P0 = -np.exp(v0)
            P0R, P0Z = P0 * v0R, P0 * v0Z
            Pv = P0 * S
            PR = P0 * acc["R"] + P0R * S
            PZ = P0 * acc["Z"] + P0Z * S
            Pp = P0 * acc["p"]
            PRR = P0 * acc["RR"] + 2 * P0R * acc["R"] + P0 * (v0RR + v0R ** 2) * S
            PZZ = P0 * acc["ZZ"] + 2 * P0Z * acc["Z"] + P0 * (v0ZZ + v0Z ** 2) * S
            PRZ = P0 * acc["RZ"] + P0R * acc["Z"] + P0Z * acc["R"] + P0 * (v0RZ + v0R * v0Z) * S
            PRp = P0 * acc["Rp"] + P0R * acc["p"]
            PZp = P0 * acc["Zp"] + P0Z * acc["p"]
            Ppp = P0 * acc["pp"]
        else:
            Pv, PR, PZ, Pp = S, acc["R"], acc["Z"], acc["p"]
            PRR, PZZ, PRZ, PRp, PZp, Ppp = acc["RR"], acc["ZZ"], acc["RZ"], acc["Rp"], acc["Zp"], acc["pp"]
        Rs = self.Rscale
        jR = 1 / np.sqrt(R * R + Rs * Rs)
        jz = 1 / np.sqrt(x[:, 2] ** 2 + Rs * Rs)
        dR, dZ = PR * jR, PZ * jz
        HRR = PRR * jR ** 2 - PR * R * jR ** 3
        HZZ = PZZ * jz ** 2 - PZ * x[:, 2] * jz ** 3
        HRZ = PRZ * jR * jz
        HRp, HZp = PRp * jR, PZp * jz
        c, s = np.cos(phi_ang), np.sin(phi_ang)
        Rsafe = np.where(R > 1e-30, R, 1.0)
        gx = np.where(R > 1e-30, dR * c - Pp * s / Rsafe, 0.0)
        gy = np.where(R > 1e-30, dR * s + Pp * c / Rsafe, 0.0)
        g = np.stack([gx, gy, dZ], axis=1)
        iR_, iR2 = 1 / Rsafe, 1 / Rsafe ** 2
        hxx = HRR * c * c - 2 * HRp * c * s * iR_ + dR * iR_ * s * s + 2 * Pp * iR2 * c * s + Ppp * s * s * iR2
        hyy = HRR * s * s + 2 * HRp * c * s * iR_ + dR * iR_ * c * c - 2 * Pp * iR2 * c * s + Ppp * c * c * iR2
        hxy = HRR * c * s + HRp * (c * c - s * s) * iR_ - dR * iR_ * c * s - Pp * (c * c - s * s) * iR2 - Ppp * c * s * iR2
        hyz = HRZ * s + HZp * c * iR_
        hxz = HRZ * c - HZp * s * iR_
        # On the z axis every harmonic's spline has zero slope in lR, so
        # Phi = Phi(axis) + lR^2/2 sum_m Q_m T_m(phi) + O(lR^3) with Q_m the
        # Fourier coefficients of d2Phi/dlR2; the m = 0, +2, -2 terms are
        # (x^2 + y^2, x^2 - y^2, 2xy) Q_m / (2 Rscale^2) to second order, the
        # others depend on the direction of approach and average out over it.
        # (The reference's R = 0 branch, _cylspl_potential_kernel.cu:855-863,
        # puts the whole d2Phi/dR2 at phi = 0 on both diagonal entries.)
        on = R <= 1e-30
        if self.log_scaling:
            Pa, PaR = P0, P0R
            PaRR = P0 * (v0RR + v0R ** 2)
            Q0 = PaRR
        else:
            Pa, PaR, PaRR = 1.0, 0.0, 0.0
            Q0 = v0RR
        Qm = {}
        for m in (2, -2):
            if m in q2:
                v2, v2R, v2RR = q2[m]
                Qm[m] = Pa * v2RR + 2 * PaR * v2R + PaRR * v2
            else:
                Qm[m] = 0.0
        j2 = jR ** 2
        hxx = np.where(on, j2 * (Q0 + Qm[2]), hxx)
        hyy = np.where(on, j2 * (Q0 - Qm[2]), hyy)
        hxy = np.where(on, j2 * Qm[-2], hxy)
        hyz = np.where(on, 0.0, hyz)
        hxz = np.where(on, 0.0, hxz)
        return Pv, g, np.stack([hxx, hyy, HZZ, hxy, hyz, hxz], axis=1)
