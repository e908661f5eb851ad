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


CYL_PATTERNS = ("even", "cos", "full", "subset")


def random_cylspline(seed: int):
    """(CylSplineCoefs, description) of a random CylSpline table: mmax 0..6
    with even, cos-only, all or a random subset of harmonics, 6..32 R nodes
    (R = 0 first, then geometric), and either a z >= 0 half grid (mirrored by
    the build) or a full grid with a z-asymmetric field.  The m = 0 harmonic
    is a Miyamoto-Nagai potential, the others eps_m (-GM) R^|m| /
    (R^2 + z^2 + c^2)^((|m|+1)/2) (regular on the axis), times
    (1 + k z / sqrt(z^2 + c^2)) on full grids."""
    from agama_gpu_amd.coefs import CylSplineCoefs
    rng = np.random.default_rng(5000 + seed)
    mmax = int(rng.choice([0, 1, 2, 3, 4, 6]))
    pattern = CYL_PATTERNS[seed % len(CYL_PATTERNS)]
    half = bool(seed % 2 == 0)
    nR = int(rng.integers(6, 33))
    nzh = int(rng.integers(5, 25))
    R_max = 10.0 ** rng.uniform(0.5, 2.5)
    z_max = R_max * 10.0 ** rng.uniform(-0.5, 0.5)
    R = np.concatenate([[0.0], R_max * np.geomspace(10.0 ** -rng.uniform(1.5, 3.0), 1.0, nR - 1)])
    zp = z_max * np.geomspace(10.0 ** -rng.uniform(1.5, 3.0), 1.0, nzh)
    z = np.concatenate([[0.0], zp]) if half else np.concatenate([-zp[::-1], [0.0], zp])
    GM = 4.30091727067736e-06 * 10.0 ** rng.uniform(9.5, 11.5)
    a, b, c = (R_max * 10.0 ** rng.uniform(-2.0, -0.5) for _ in range(3))
    kz = 0.0 if half else rng.uniform(-0.5, 0.5)
    RR, ZZ = np.meshgrid(R, z, indexing="ij")
    ms = [m for m in range(-mmax, mmax + 1)]
    if pattern == "even":
        ms = [m for m in ms if m % 2 == 0]
    elif pattern == "cos":
        ms = [m for m in ms if m >= 0]
    elif pattern == "subset":
        ms = [m for m in ms if m == 0 or rng.random() < 0.5]
    phi = {}
    for m in ms:
        if m == 0:
            t = -GM / np.sqrt(RR ** 2 + (a + np.sqrt(ZZ ** 2 + b ** 2)) ** 2)
        else:
            am = abs(m)
            eps = np.clip(rng.normal(), -2.0, 2.0) * 0.1 * 0.7 ** am
            t = eps * -GM * RR ** am / (RR ** 2 + ZZ ** 2 + c ** 2) ** ((am + 1) / 2)
            t = t * (1.0 + kz * ZZ / np.sqrt(ZZ ** 2 + c ** 2))
        phi[m] = np.asarray(t, float)
    coefs = CylSplineCoefs(R, z, phi, mmax)
    desc = dict(seed=seed, mmax=mmax, pattern=pattern, half=half, nR=nR, nz=len(z),
                R_max=float(R_max), z_max=float(z_max), ms=ms)
    return coefs, desc


def cyl_sweep_points(desc: dict, n: int, seed: int) -> np.ndarray:
    """Radii log-uniform over [R_max / 1000, 5 max(R_max, z_max)], isotropic
    directions (inside and outside the grid box); 2 % within 1e-9 R_max of
    the z axis, a few on it and in the midplane."""
    rng = np.random.default_rng(seed)
    top = 5.0 * max(desc["R_max"], desc["z_max"])
    r = np.exp(rng.uniform(np.log(desc["R_max"] / 1000.0), np.log(top), n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    x = u * r[:, None]
    k = max(1, n // 50)
    x[:k, :2] *= (1e-9 * desc["R_max"] / np.maximum(np.hypot(x[:k, 0], x[:k, 1]), 1e-300))[:, None]
    x[k:k + 3] = [[0.0, 0.0, 0.3 * desc["z_max"]], [0.0, 0.0, -0.6 * desc["z_max"]],
                  [0.4 * desc["R_max"], 0.2 * desc["R_max"], 0.0]]
    return x
