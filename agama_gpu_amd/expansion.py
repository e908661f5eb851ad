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
quadrature; dPhi_lm/dr follows from the same two integrals.
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


def _cumtrapz_log(f: np.ndarray, lnr: np.ndarray) -> np.ndarray:
    """Cumulative integral of f d(ln r) by the composite Simpson-corrected
    trapezoid on a uniform ln r grid (f sampled on the grid)."""
    h = lnr[1] - lnr[0]
    out = np.zeros_like(f)
    # trapezoid with end-point derivative correction (4th order on smooth f)
    out[1:] = np.cumsum(0.5 * h * (f[1:] + f[:-1]), axis=0)
    return out


def multipole_from_density(rho: Callable[[np.ndarray, np.ndarray, np.ndarray], np.ndarray],
                           lmax: int, r_min: float, r_max: float, n_r: int,
                           symmetry: str = "none", n_theta: Optional[int] = None,
                           n_phi: Optional[int] = None, fine_factor: int = 40,
                           r_in: Optional[float] = None, r_out: Optional[float] = None
                           ) -> MultipoleCoefs:
    """Solve Poisson for an arbitrary density rho(x, y, z) [Msol/kpc^3].

    ``symmetry`` -- "none", "triaxial" (even l, even m >= 0 only) or
    "axisymmetric" (m = 0, even l): harmonics forbidden by the symmetry are
    set to exactly zero.
    """
    n_theta = n_theta or (lmax + 24)
    n_phi = n_phi or (2 * lmax + 24)
    xg, wg = np.polynomial.legendre.leggauss(n_theta)
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

    # fine log grid that contains the output nodes
    lo = math.log(r_in if r_in else r_min * 1e-4)
    hi = math.log(r_out if r_out else r_max * 1e3)
    node_ln = np.linspace(math.log(r_min), math.log(r_max), n_r)
    dl = (node_ln[1] - node_ln[0]) / fine_factor
    n_lo = int(math.ceil((node_ln[0] - lo) / dl))
    n_hi = int(math.ceil((hi - node_ln[-1]) / dl))
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
        I_in = _cumtrapz_log(f_in, lnr)
        # inner tail below the fine grid: rho ~ power law; include its integral
        I_out_cum = _cumtrapz_log(f_out, lnr)
        I_out = I_out_cum[-1] - I_out_cum
        rn = r[node_idx]
        a_in, a_out = I_in[node_idx], I_out[node_idx]
        k = -4.0 * math.pi * G_AGAMA / (2 * l + 1)
        phi[:, j] = k * (rn ** (-(l + 1)) * a_in + rn ** l * a_out)
        dphi[:, j] = k * (-(l + 1) * rn ** (-(l + 2)) * a_in + l * rn ** (l - 1) * a_out)
    return MultipoleCoefs(np.exp(node_ln), phi, dphi, labels, lmax,
                          symmetry={"none": "None", "triaxial": "Triaxial",
                                    "axisymmetric": "Axisymmetric"}.get(symmetry, "None"))


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
