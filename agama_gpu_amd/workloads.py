"""Synthetic workloads: seeded positions and coefficient tables.

The measured workload is a Multipole halo with lmax=8 on 50 log-spaced radial
nodes over r in [0.1, 500] kpc, built from a triaxial Dehnen density
(gamma=1, M=1e12 Msol, a=20 kpc, axis ratios 1:0.8:0.6), evaluated for
potential+force at positions drawn from an isotropic Hernquist profile
(a=20 kpc) with 1% of the points log-uniform in r in [0.01, 2000] kpc so both
extrapolation branches run.  The point-drawing call order below is frozen.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from .coefs import MultipoleCoefs
from .expansion import dehnen_density, multipole_from_density


def hernquist_points(n: int, seed: int = 0, a: float = 20.0, frac_loguniform: float = 0.01,
                     r_lo: float = 0.01, r_hi: float = 2000.0) -> np.ndarray:
    """(n,3) float64 positions: Hernquist radii r = a sqrt(u)/(1-sqrt(u)),
    isotropic directions, then a fraction replaced by log-uniform radii."""
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    su = np.sqrt(u)
    r = a * su / (1.0 - su)
    ct = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    n_ext = int(n * frac_loguniform)
    if n_ext:
        idx = rng.choice(n, n_ext, replace=False)
        r[idx] = 10.0 ** rng.uniform(math.log10(r_lo), math.log10(r_hi), n_ext)
    st = np.sqrt(1.0 - ct * ct)
    return np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)


def loguniform_points(n: int, seed: int, r_lo: float, r_hi: float) -> np.ndarray:
    """Isotropic directions, radii uniform in log r."""
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(math.log(r_lo), math.log(r_hi), n))
    ct = rng.uniform(-1.0, 1.0, n)
    ph = rng.uniform(0.0, 2.0 * math.pi, n)
    st = np.sqrt(1.0 - ct * ct)
    return np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)


@lru_cache(maxsize=4)
def halo_lmax8_coefs() -> MultipoleCoefs:
    """Triaxial Dehnen (gamma=1, M=1e12, a=20, 1:0.8:0.6) -> lmax 8, 50 nodes
    on [0.1, 500] kpc; harmonics forbidden by triaxial symmetry are exactly 0."""
    rho = dehnen_density(1e12, 20.0, 1.0, 0.8, 0.6)
    return multipole_from_density(rho, 8, 0.1, 500.0, 50, symmetry="triaxial")
