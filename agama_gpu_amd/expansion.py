"""Multipole coefficient builder: density -> Agama Multipole table (CPU).

The reference obtains every table from Agama's C++ Multipole constructor
(``agama.Potential(type='Multipole', density=...)`` then ``export``;
reference gpu_potential.py:688-727, 1935-2052, tests/test_phase1_multipole.py
:107-129).  Agama is not available here, so this module solves Poisson's
equation for the spherical-harmonic expansion itself and emits a
``MultipoleCoefs`` in Agama's convention:

    Phi(r, theta, phi) = sum_lm Phi_lm(r) * B_lm(theta, phi),
    B_lm = mul_m * Ptilde_lm(cos theta) * {cos(m phi) for m >= 0, sin(|m| phi) for m < 0},
    Ptilde_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m  (Condon-Shortley phase),
    mul_0 = 2 sqrt(pi), mul_m = 2 sqrt(2 pi)

so that (1/4pi) * integral B_lm B_l'm' dOmega = delta (B_00 = 1).  For each
harmonic the radial solution is the Green's-function integral

    Phi_lm(r) = -4 pi G / (2l+1) * [ r^-(l+1) int_0^r rho_lm r'^(l+2) dr'
                                     + r^l int_r^inf rho_lm r'^(1-l) dr' ]

with rho_lm(r) = (1/4pi) integral rho B_lm dOmega by Gauss-Legendre x uniform-phi
quadrature (Gauss-Legendre on each hemisphere separately with ``split_equator``,
for densities with a kink at the midplane); dPhi_lm/dr follows from the same
two integrals.  The radial integrals run over a fine uniform grid in ln r
(``fine_factor`` steps per node interval, from 1e-4 r_min to 1e3 r_max by
default) by the composite Simpson rule, plus the parts below and above that
grid integrated as power laws fitted to its two end samples.

Builders on top of it, standing in for the types the reference obtains from
Agama (gpu_potential.py:1935-2052): ``spheroid_coefs`` (Agama's Spheroid
density: double power law with an optional exponential cutoff, triaxial),
``disk_potential_parts`` (DiskAnsatz plus an axisymmetric Multipole of the
residual density, Agama's construction of its Disk potential) and
``king_coefs`` (the lowered isothermal sphere, its potential from the Poisson
ODE).
"""
from __future__ import annotations

import math
from typing import Callable, List, Optional, Tuple

import numpy as np

from .coefs import MultipoleCoefs
from .multipole_prep import G_AGAMA


def lm_labels(lmax: int) -> List[Tuple[int, int]]:
    """Agama column order: l-major, m = -l..l."""
    return [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]


def real_basis(lmax: int, ct: np.ndarray, ph: np.ndarray) -> np.ndarray:
    """B_lm at the given (cos theta, phi) points; shape (n_lm, npts)."""
    from scipy.special import sph_harm_y
    th = np.arccos(np.clip(ct, -1.0, 1.0))
    out = []
    for l, m in lm_labels(lmax):
        am = abs(m)
        Y = sph_harm_y(l, am, th, ph)                  # Ptilde_lm e^{i m phi}
        Pt = np.real(Y * np.exp(-1j * am * ph))
        mul = 2.0 * math.sqrt(math.pi) if m == 0 else 2.0 * math.sqrt(2.0 * math.pi)
        T = np.cos(am * ph) if m >= 0 else np.sin(am * ph)
        out.append(mul * Pt * T)
    return np.asarray(out)


def _cum_simpson(f: np.ndarray, h: float, from_end: bool = False) -> np.ndarray:
    """Cumulative integral of f (rows: a uniform grid of spacing h, an odd
    number of them) by the composite Simpson rule (4th order in h), at every
    even-numbered row: from the first row to it, or with ``from_end`` from it
    to the last row (accumulated from that end, so that a large integrand at
    the other end does not swamp it by cancellation); odd rows are NaN."""
    n = f.shape[0]
    if n % 2 == 0:
        raise ValueError("Simpson grid needs an odd number of points")
    out = np.full_like(f, np.nan)
    panels = (h / 3.0) * (f[0:-2:2] + 4.0 * f[1:-1:2] + f[2::2])
    if from_end:
        out[-1] = 0.0
        out[0:-2:2] = np.cumsum(panels[::-1], axis=0)[::-1]
    else:
        out[0] = 0.0
        out[2::2] = np.cumsum(panels, axis=0)
    return out


def _power_tail(f0, f1, dl: float, outward: bool) -> float:
    """Integral over ln r of a power-law tail f = f0 exp(p (ln r - ln r0))
    beyond an end r0 of the grid (outward: from r0 to infinity, needs p < 0;
    else from 0 to r0, needs p > 0), with p from the end sample f0 and the
    sample f1 one step dl inwards.  0 without a usable power law (a zero or a
    sign change, or a slope of the wrong sign or below 1/2 in magnitude)."""
    f0, f1 = float(f0), float(f1)
    if f0 == 0.0 or f0 * f1 <= 0.0:
        return 0.0
    p = math.log(f0 / f1) / dl * (1.0 if outward else -1.0)
    if outward:
        return -f0 / p if p < -0.5 else 0.0
    return f0 / p if p > 0.5 else 0.0


def multipole_from_density(rho: Callable[[np.ndarray, np.ndarray, np.ndarray], np.ndarray],
                           lmax: int, r_min: float, r_max: float, n_r: int,
                           symmetry: str = "none", n_theta: Optional[int] = None,
                           n_phi: Optional[int] = None, fine_factor: int = 40,
                           r_in: Optional[float] = None, r_out: Optional[float] = None,
                           split_equator: bool = False) -> MultipoleCoefs:
    """Solve Poisson for an arbitrary density rho(x, y, z) [Msol/kpc^3].

    ``symmetry`` -- "none", "triaxial" (even l, even m >= 0 only),
    "axisymmetric" (m = 0, even l) or "spherical" (l = 0 only): harmonics
    forbidden by the symmetry are set to exactly zero (and, for the last
    two, the density is sampled at one azimuth).  ``split_equator``: n_theta
    Gauss-Legendre nodes on each hemisphere instead of n_theta over the
    sphere, for densities with a kink at z = 0.
    """
    n_theta = n_theta or (lmax + 24)
    n_phi = n_phi or (2 * lmax + 24)
    if symmetry in ("axisymmetric", "spherical"):
        n_phi = 1
    xg, wg = np.polynomial.legendre.leggauss(n_theta)
    if split_equator:
        xg, wg = np.concatenate([0.5 * (xg - 1.0), 0.5 * (xg + 1.0)]), np.concatenate([wg, wg]) * 0.5
    phis = (np.arange(n_phi) + 0.5) * (2.0 * math.pi / n_phi)
    CT, PH = np.meshgrid(xg, phis, indexing="ij")
    W = np.repeat(wg[:, None], n_phi, axis=1) * (2.0 * math.pi / n_phi) / (4.0 * math.pi)
    ct, ph, w = CT.ravel(), PH.ravel(), W.ravel()
    st = np.sqrt(1.0 - ct * ct)
    B = real_basis(lmax, ct, ph)                                   # (n_lm, nang)
    labels = lm_labels(lmax)
    allowed = np.ones(len(labels), bool)
    for j, (l, m) in enumerate(labels):
        if symmetry == "triaxial":
            allowed[j] = (l % 2 == 0) and (m % 2 == 0) and m >= 0
        elif symmetry == "axisymmetric":
            allowed[j] = (m == 0) and (l % 2 == 0)
        elif symmetry == "spherical":
            allowed[j] = l == 0
        elif symmetry != "none":
            raise ValueError(f"unknown symmetry {symmetry!r}")

    # fine log grid that contains the output nodes at even positions (the
    # Simpson rule gives cumulative integrals there), with an odd length
    if fine_factor < 2 or fine_factor % 2:
        raise ValueError("fine_factor must be even and >= 2")
    lo = math.log(r_in if r_in else r_min * 1e-4)
    hi = math.log(r_out if r_out else r_max * 1e3)
    node_ln = np.linspace(math.log(r_min), math.log(r_max), n_r)
    dl = (node_ln[1] - node_ln[0]) / fine_factor
    n_lo = 2 * int(math.ceil((node_ln[0] - lo) / (2 * dl)))
    n_hi = 2 * int(math.ceil((hi - node_ln[-1]) / (2 * dl)))
    lnr = node_ln[0] + dl * np.arange(-n_lo, (n_r - 1) * fine_factor + n_hi + 1)
    r = np.exp(lnr)
    node_idx = n_lo + fine_factor * np.arange(n_r)

    rho_lm = np.empty((len(r), len(labels)))
    chunk = 256
    for i0 in range(0, len(r), chunk):
        rr = r[i0:i0 + chunk, None]
        x = rr * (st * np.cos(ph))[None, :]
        y = rr * (st * np.sin(ph))[None, :]
        z = rr * ct[None, :]
        vals = rho(x, y, z)                                        # (chunk, nang)
        rho_lm[i0:i0 + chunk] = (vals * w[None, :]) @ B.T
    rho_lm[:, ~allowed] = 0.0

    phi = np.zeros((n_r, len(labels)))
    dphi = np.zeros((n_r, len(labels)))
    for j, (l, m) in enumerate(labels):
        if not allowed[j]:
            continue
        f_in = rho_lm[:, j] * r ** (l + 3)            # d/dlnr of int rho r^(l+2) dr
        f_out = rho_lm[:, j] * r ** (2 - l)           # d/dlnr of int rho r^(1-l) dr
        # the parts below and above the fine grid as power-law tails
        # (rho ~ r^-gamma inside, ~ r^-beta outside)
        I_in = _cum_simpson(f_in, dl) + _power_tail(f_in[0], f_in[1], dl, outward=False)
        I_out = _cum_simpson(f_out, dl, from_end=True) + _power_tail(f_out[-1], f_out[-2], dl,
                                                                      outward=True)
        rn = r[node_idx]
        a_in, a_out = I_in[node_idx], I_out[node_idx]
        k = -4.0 * math.pi * G_AGAMA / (2 * l + 1)
        phi[:, j] = k * (rn ** (-(l + 1)) * a_in + rn ** l * a_out)
        dphi[:, j] = k * (-(l + 1) * rn ** (-(l + 2)) * a_in + l * rn ** (l - 1) * a_out)
    return MultipoleCoefs(np.exp(node_ln), phi, dphi, labels, lmax,
                          symmetry={"none": "None", "triaxial": "Triaxial",
                                    "axisymmetric": "Axisymmetric",
                                    "spherical": "Spherical"}.get(symmetry, "None"))


def dehnen_density(mass: float, a: float, gamma: float, p: float = 1.0, q: float = 1.0):
    """Triaxial Dehnen density with axis ratios (1, p, q)."""
    norm = (3.0 - gamma) * mass / (4.0 * math.pi * a ** 3 * p * q)

    def rho(x, y, z):
        m = np.sqrt(x * x + (y / p) ** 2 + (z / q) ** 2) / a
        m = np.maximum(m, 1e-300)
        return norm * m ** (-gamma) * (1.0 + m) ** (gamma - 4.0)
    return rho


def multipole_from_potential(phi_fn: Callable, dphi_fn: Callable, r_min: float, r_max: float,
                             n_r: int) -> MultipoleCoefs:
    """Monopole-only table from a closed-form spherical potential Phi(r), dPhi/dr."""
    r = np.exp(np.linspace(math.log(r_min), math.log(r_max), n_r))
    return MultipoleCoefs(r, phi_fn(r)[:, None], dphi_fn(r)[:, None], [(0, 0)], 0, "Spherical")


# ---- Agama's Spheroid ----------------------------------------------------------
def spheroid_density(densityNorm: float = 1.0, scaleRadius: float = 1.0, alpha: float = 1.0,
                     beta: float = 4.0, gamma: float = 1.0, outerCutoffRadius: float = 0.0,
                     cutoffStrength: float = 2.0, axisRatioY: float = 1.0,
                     axisRatioZ: float = 1.0):
    """rho = rho0 (m/a)^-gamma [1 + (m/a)^alpha]^((gamma-beta)/alpha) exp(-(m/r_cut)^xi),
    m = sqrt(x^2 + (y/p)^2 + (z/q)^2): the density of Agama's Spheroid type
    (no cutoff when outerCutoffRadius is 0)."""
    rho0, a, p, q = float(densityNorm), float(scaleRadius), float(axisRatioY), float(axisRatioZ)
    al, be, ga = float(alpha), float(beta), float(gamma)
    rc, xi = float(outerCutoffRadius), float(cutoffStrength)

    def rho(x, y, z):
        m = np.sqrt(x * x + (y / p) ** 2 + (z / q) ** 2) / a
        m = np.where(m < 1e-300, 1e-300, m)
        v = rho0 * m ** (-ga) * (1.0 + m ** al) ** ((ga - be) / al)
        if rc > 0.0:
            v = v * np.exp(-(m * a / rc) ** xi)
        return v
    return rho


def spheroid_mass(densityNorm=1.0, scaleRadius=1.0, alpha=1.0, beta=4.0, gamma=1.0,
                  outerCutoffRadius=0.0, cutoffStrength=2.0, axisRatioY=1.0, axisRatioZ=1.0):
    """Total mass of the Spheroid density: 4 pi p q rho0 a^3 int s^(2-gamma)
    (1+s^alpha)^((gamma-beta)/alpha) exp(-(s a/r_cut)^xi) ds; without a cutoff
    that is B((3-gamma)/alpha, (beta-3)/alpha) / alpha (finite for beta > 3)."""
    from scipy import integrate, special
    a, al, be, ga = float(scaleRadius), float(alpha), float(beta), float(gamma)
    pre = 4.0 * math.pi * float(axisRatioY) * float(axisRatioZ) * float(densityNorm) * a ** 3
    if ga >= 3.0:
        raise ValueError("Spheroid: gamma must be below 3 for a finite mass")
    if outerCutoffRadius <= 0.0:
        if be <= 3.0:
            raise ValueError("Spheroid: infinite mass for beta <= 3 without an outer cutoff")
        return pre * special.beta((3.0 - ga) / al, (be - 3.0) / al) / al
    rc, xi = float(outerCutoffRadius), float(cutoffStrength)

    def f(t):       # s = e^t
        s = math.exp(t)
        return s ** (3.0 - ga) * (1.0 + s ** al) ** ((ga - be) / al) * math.exp(-(s * a / rc) ** xi)
    hi = math.log(rc / a) + math.log(60.0) / xi
    val, _ = integrate.quad(f, -60.0, hi, limit=400, epsabs=0.0, epsrel=1e-13)
    return pre * val


def spheroid_coefs(lmax: int = 8, gridSizeR: int = 50, rmin: Optional[float] = None,
                   rmax: Optional[float] = None, mass: Optional[float] = None, **kw
                   ) -> MultipoleCoefs:
    """Multipole table of Agama's Spheroid (keywords as spheroid_density,
    ``mass`` in place of ``densityNorm`` normalises the total mass): spherical
    for equal axes (l = 0 only), triaxial symmetry otherwise."""
    kw = dict(kw)
    if mass is not None:
        kw["densityNorm"] = 1.0
        kw["densityNorm"] = float(mass) / spheroid_mass(**kw)
    a = float(kw.get("scaleRadius", 1.0))
    p, q = float(kw.get("axisRatioY", 1.0)), float(kw.get("axisRatioZ", 1.0))
    sym = "spherical" if (abs(p - 1.0) < 1e-10 and abs(q - 1.0) < 1e-10) else "triaxial"
    rmin = rmin if rmin is not None else 1e-3 * a
    rmax = rmax if rmax is not None else 1e3 * a
    if sym == "spherical":
        lmax = 0
    return multipole_from_density(spheroid_density(**kw), int(lmax), rmin, rmax, int(gridSizeR),
                                  symmetry=sym)


# ---- Agama's Disk: DiskAnsatz + Multipole of the residual -----------------------
def _vertical(z, h):
    """H(z), H'(z), H''(z) of the DiskAnsatz vertical function (analytic_dev.h):
    exponential for h > 0, isothermal for h < 0, thin (the delta function of
    H'' left out) for h = 0.  H'' is the disk's vertical density profile."""
    az, sg = np.abs(z), np.sign(z)
    if h > 0:
        u = az / h
        eu = np.exp(-u)
        return 0.5 * h * (eu - 1.0 + u), 0.5 * sg * (1.0 - eu), 0.5 * eu / h
    if h < 0:
        hh = -h
        u = az / hh
        eu = np.exp(-2.0 * u)
        th = (1.0 - eu) / (1.0 + eu)
        return 0.5 * hh * (u + np.log1p(eu) - math.log(2.0)), 0.5 * sg * th, 0.5 * (1.0 - th * th) / hh
    return 0.5 * az, 0.5 * sg, 0.0 * az


def disk_residual_density(surfaceDensity: float, scaleRadius: float, scaleHeight: float,
                          innerCutoffRadius: float = 0.0):
    """rho_disk - rho_ansatz with rho_disk = Sigma exp(-R/R_d - R_in/R) H''(z)
    and rho_ansatz = Laplacian(Phi_ansatz) / (4 pi G), Phi_ansatz =
    4 pi G Sigma exp(-r/R_d - R_in/r) H(z) (r spherical): the density whose
    Multipole expansion completes the DiskAnsatz to the disk's potential
    (reference gpu_potential.py:1984-2000).  The delta-function parts of a thin
    disk cancel and are left out of both."""
    S, Rd, h, rin = float(surfaceDensity), float(scaleRadius), float(scaleHeight), \
        float(innerCutoffRadius)

    def rho(x, y, z):
        R = np.sqrt(x * x + y * y)
        r = np.sqrt(R * R + z * z)
        r = np.where(r < 1e-300, 1e-300, r)
        Hv, H1, H2 = _vertical(z, h)
        with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
            fR = np.where(R > 0, np.exp(-R / Rd - (rin / np.where(R > 0, R, 1.0) if rin else 0.0)),
                          0.0 if rin else 1.0)
        ir = 1.0 / r
        f = S * np.exp(-r / Rd - rin * ir)
        fo = -1.0 / Rd + rin * ir * ir
        f1, f2 = f * fo, f * (fo * fo - 2.0 * rin * ir ** 3)
        ansatz = Hv * (f2 + 2.0 * f1 * ir) + 2.0 * f1 * H1 * z * ir + f * H2
        return S * fR * H2 - ansatz
    return rho


def disk_potential_parts(surfaceDensity: Optional[float] = None, scaleRadius: float = 1.0,
                         scaleHeight: float = 0.1, innerCutoffRadius: float = 0.0,
                         mass: Optional[float] = None, sersicIndex: float = 1.0,
                         lmax: int = 32, gridSizeR: int = 50, rmin: Optional[float] = None,
                         rmax: Optional[float] = None, n_theta: int = 48):
    """(DiskAnsatz parameters, Multipole table of the residual density) of an
    exponential disk -- the two components of Agama's Disk potential
    (reference gpu_potential.py:1984-2052).  ``mass`` in place of
    ``surfaceDensity``: M = 2 pi Sigma int R exp(-R/R_d - R_in/R) dR."""
    from scipy import integrate
    if abs(float(sersicIndex) - 1.0) > 1e-12:
        raise NotImplementedError("Disk: only sersicIndex = 1 (the exponential disk) is supported")
    Rd, h, rin = float(scaleRadius), float(scaleHeight), float(innerCutoffRadius)
    if mass is not None:
        if rin > 0.0:
            val = integrate.quad(lambda R: R * math.exp(-R / Rd - rin / R), 0.0, math.inf,
                                 epsabs=0.0, epsrel=1e-13, limit=400)[0]
        else:
            val = Rd * Rd
        surfaceDensity = float(mass) / (2.0 * math.pi * val)
    if surfaceDensity is None:
        surfaceDensity = 1.0
    S = float(surfaceDensity)
    rmin = rmin if rmin is not None else 1e-2 * min(Rd, abs(h)) if h else 1e-2 * Rd
    rmax = rmax if rmax is not None else 200.0 * Rd
    coefs = multipole_from_density(disk_residual_density(S, Rd, h, rin), int(lmax), rmin, rmax,
                                   int(gridSizeR), symmetry="axisymmetric", n_theta=n_theta,
                                   split_equator=True, fine_factor=20)
    ansatz = dict(surfaceDensity=S, scaleRadius=Rd, scaleHeight=h, innerCutoffRadius=rin)
    return ansatz, coefs


# ---- King (lowered isothermal) sphere ------------------------------------------
def king_profile(W0: float, trunc: float = 1.0):
    """Dimensionless King model: W(xi) from W'' + 2 W'/xi = -9 rho(W)/rho(W0),
    W(0) = W0, W'(0) = 0, with rho(W) = e^W P(g + 3/2, W) (P the regularised
    lower incomplete gamma function, g = ``trunc``; g = 1 is King's 1966
    model), xi = r / r_0 with the King radius r_0^2 = 9 sigma^2 / (4 pi G rho_0).
    Returns (xi, W, dW/dxi) on the solver's dense output up to the tidal radius
    xi_t where W = 0, and xi_t."""
    from scipy import integrate, special
    g = float(trunc)
    W0 = float(W0)

    def rho_hat(W):
        W = np.maximum(W, 0.0)
        return np.exp(W) * special.gammainc(g + 1.5, W)
    rho0 = float(rho_hat(np.array(W0)))

    def rhs(xi, y):
        W, dW = y
        return [dW, -9.0 * float(rho_hat(np.array(W))) / rho0 - 2.0 * dW / xi]

    def edge(xi, y):
        return y[0]
    edge.terminal, edge.direction = True, -1
    xi0 = 1e-6                                  # series start: W = W0 - 3/2 xi^2
    sol = integrate.solve_ivp(rhs, (xi0, 1e8), [W0 - 1.5 * xi0 ** 2, -3.0 * xi0], method="DOP853",
                              rtol=1e-13, atol=1e-15, events=edge, dense_output=True)
    if not sol.t_events[0].size:
        raise ValueError(f"King W0={W0}: no tidal radius reached")
    return sol, float(sol.t_events[0][0])


def king_coefs(mass: float = 1.0, scaleRadius: float = 1.0, W0: float = 7.0, trunc: float = 1.0,
               gridSizeR: int = 60) -> MultipoleCoefs:
    """Monopole Multipole table of a King model of total ``mass`` and King
    radius ``scaleRadius``: Phi = Phi(r_t) - sigma^2 W(r / r_0) inside the tidal
    radius r_t, -G M / r outside, with sigma^2 = G M / (r_0 xi_t^2 |W'(xi_t)|)
    (continuity of the force at r_t); nodes log-spaced from 1e-3 r_0 to 100 r_t."""
    sol, xi_t = king_profile(W0, trunc)
    r0 = float(scaleRadius)
    GM = G_AGAMA * float(mass)
    dWt = float(sol.sol(xi_t)[1])
    sig2 = GM / (r0 * xi_t ** 2 * abs(dWt))
    rt = xi_t * r0
    r = np.exp(np.linspace(math.log(1e-3 * r0), math.log(100.0 * rt), int(gridSizeR)))
    phi = np.empty_like(r)
    dphi = np.empty_like(r)
    ins = r < rt
    if np.any(ins):
        W, dW = sol.sol(r[ins] / r0)
        phi[ins] = -GM / rt - sig2 * W
        dphi[ins] = -sig2 * dW / r0
    phi[~ins] = -GM / r[~ins]
    dphi[~ins] = GM / r[~ins] ** 2
    return MultipoleCoefs(r, phi[:, None], dphi[:, None], [(0, 0)], 0, "Spherical")
