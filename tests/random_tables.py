"""Seeded random Multipole tables for parity sweeps (test data, not a model).

random_multipole(seed) draws a table whose structure varies with the seed:
lmax 0..12, 6..60 radial nodes on a random log-uniform range, and one of
several harmonic patterns (every (l, m) with sin terms, triaxial even-even,
axisymmetric, even-l only, odd-m sin-only mixtures, a random subset).  The
monopole is a Hernquist-like -GM/(r+a), optionally lifted by a constant so
that it changes sign (no log-scaling); the l > 0 columns are
eps_lm * (-GM r^l / (r+a)^(l+1)) with decaying random amplitudes, and every
dPhi/dr column is the closed-form derivative, so the file is self-consistent.
"""
from __future__ import annotations

import numpy as np

from agama_gpu_amd.coefs import MultipoleCoefs

PATTERNS = ("full", "triaxial", "axisym", "even_l", "sin_odd", "subset")


def _pattern_keep(pattern: str, rng):
    if pattern == "full":
        return lambda l, m: True
    if pattern == "triaxial":
        return lambda l, m: l % 2 == 0 and m >= 0 and m % 2 == 0
    if pattern == "axisym":
        return lambda l, m: m == 0
    if pattern == "even_l":
        return lambda l, m: l % 2 == 0
    if pattern == "sin_odd":
        # m = 0 plus sin terms of odd |m|: programs with sin-only groups
        return lambda l, m: m == 0 or (m < 0 and (-m) % 2 == 1)
    chosen = {}

    def keep(l, m):
        if (l, m) == (0, 0):
            return True
        if (l, m) not in chosen:
            chosen[(l, m)] = bool(rng.random() < 0.45)
        return chosen[(l, m)]
    return keep


def random_multipole(seed: int, lmax: int = None, pattern: str = None, signed: bool = None):
    """(MultipoleCoefs, description) of a random self-consistent table."""
    rng = np.random.default_rng(1000 + seed)
    if lmax is None:
        lmax = int(rng.integers(0, 13))
    if pattern is None:
        pattern = PATTERNS[seed % len(PATTERNS)]
    if signed is None:
        signed = bool(seed % 5 == 3)
    nR = int(rng.integers(6, 61))
    r_min = 10.0 ** rng.uniform(-2.0, 0.0)
    r_max = r_min * 10.0 ** rng.uniform(2.0, 4.0)
    R = r_min * np.exp(np.arange(nR) * (np.log(r_max / r_min) / (nR - 1)))
    GM = 4.30091727067736e-06 * 10.0 ** rng.uniform(10.0, 12.5)
    a = 10.0 ** rng.uniform(np.log10(r_min) + 0.5, np.log10(r_max) - 0.5)
    labels = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
    keep = _pattern_keep(pattern, rng)
    phi = np.zeros((nR, len(labels)))
    dphi = np.zeros((nR, len(labels)))
    for j, (l, m) in enumerate(labels):
        if not keep(l, m):
            continue
        if l == 0:
            phi[:, j] = -GM / (R + a)
            dphi[:, j] = GM / (R + a) ** 2
            if signed:      # positive from r = (r_max - a) / 2 out: no log-scaling
                phi[:, j] += 2.0 * GM / (r_max + a)
            continue
        eps = rng.normal() * 0.4 * 0.6 ** l
        g = -GM * R ** l / (R + a) ** (l + 1)
        dg = -GM * R ** (l - 1) * (l * a - R) / (R + a) ** (l + 2)
        phi[:, j] = eps * g
        dphi[:, j] = eps * dg
    c = MultipoleCoefs(R, phi, dphi, labels, lmax)
    desc = dict(seed=seed, lmax=lmax, pattern=pattern, signed=signed, nR=nR,
                r_min=float(r_min), r_max=float(r_max), a=float(a), GM=float(GM))
    return c, desc


def sweep_points(desc: dict, n: int, seed: int) -> np.ndarray:
    """Radii log-uniform over [r_min / 20, 20 r_max] (both extrapolation laws),
    isotropic directions; 2 % within 1e-8 r of the z axis, a few on it, at the
    origin and just inside and just outside r_max (r_max itself may take
    either branch, depending on the last bit of ln r: the edge tests cover it)."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(np.log(desc["r_min"] / 20.0), np.log(desc["r_max"] * 20.0), n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    x = u * r[:, None]
    k = max(1, n // 50)
    x[:k, 0] = r[:k] * 1e-8 * np.cos(1.0 + np.arange(k))
    x[:k, 1] = r[:k] * 1e-8 * np.sin(1.0 + np.arange(k))
    x[:k, 2] = r[:k] * np.sign(rng.normal(size=k))
    x[k:k + 5] = [[0, 0, desc["r_min"] * 3], [0, 0, -desc["r_max"] / 3], [0, 0, 0],
                  [desc["r_max"] * (1 - 1e-9), 0, 0], [0, desc["r_max"] * (1 + 1e-9), 0]]
    return x
