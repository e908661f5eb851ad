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


def ensemble_spec(n_pot: int = 256, seed: int = 3):
    """Shapes of the ragged ensemble: per potential (lmax in 4..12, radial
    nodes in 30..80, axis ratios p >= q in [0.6, 1], particle count
    log-uniform in [1e3, 1e6]).  The draw order is frozen."""
    rng = np.random.default_rng(seed)
    lmax = rng.integers(4, 13, n_pot)
    n_r = rng.integers(30, 81, n_pot)
    pq = np.sort(rng.uniform(0.6, 1.0, (n_pot, 2)), axis=1)[:, ::-1]
    counts = np.floor(np.exp(rng.uniform(math.log(1e3), math.log(1e6), n_pot))).astype(np.int64)
    return [dict(lmax=int(lmax[j]), n_r=int(n_r[j]), p=float(pq[j, 0]), q=float(pq[j, 1]),
                 count=int(counts[j])) for j in range(n_pot)]


def ensemble_coefs(spec: dict) -> MultipoleCoefs:
    """Triaxial Dehnen (gamma=1, M=1e12, a=20 kpc, axes 1:p:q) on spec's
    lmax and radial nodes over [0.1, 500] kpc.  The quadrature is coarser
    than halo_lmax8_coefs' (the ensemble measures throughput; its tables
    only need to be smooth and of the right shape)."""
    rho = dehnen_density(1e12, 20.0, 1.0, spec["p"], spec["q"])
    L = spec["lmax"]
    return multipole_from_density(rho, L, 0.1, 500.0, spec["n_r"], symmetry="triaxial",
                                  n_theta=L + 10, n_phi=2 * L + 10, fine_factor=6,
                                  r_in=1e-3, r_out=5e4)


def fit_problem(n: int = 1_000_000, n_knots: int = 200, t_eval: float = 0.73, device=None):
    """Force-fitting problem: the lmax=8 halo table (all 81 harmonics live)
    shifted along a cubic-spline centre trajectory of n_knots knots on
    t in [0, 2] (c(t) = (30 cos wt, 30 sin wt, 3 sin 2 pi t), w = 2 pi / 1.5),
    evaluated at time t_eval at n Hernquist tracers (seed 2).  The target
    forces come from a seed-2 perturbed copy of the table (every column
    scaled by 1 + 0.01 N(0,1)).  Returns (model, xyz, target_force)."""
    import copy

    import torch

    from .autograd import FittableMultipole, FittableShiftedMultipole, SplineCenter
    coefs = halo_lmax8_coefs()
    rng = np.random.default_rng(2)
    fac = 1.0 + 0.01 * rng.normal(size=coefs.phi.shape[1])
    target = copy.deepcopy(coefs)
    target.phi = np.asarray(coefs.phi) * fac[None, :]
    target.dphi_dr = np.asarray(coefs.dphi_dr) * fac[None, :]
    times = np.linspace(0.0, 2.0, n_knots)
    w = 2.0 * math.pi / 1.5
    knots = np.stack([30.0 * np.cos(w * times), 30.0 * np.sin(w * times),
                      3.0 * np.sin(2.0 * math.pi * times)], axis=1)
    xyz = torch.as_tensor(hernquist_points(n, 2), device=device)
    model = FittableShiftedMultipole(FittableMultipole(coefs, device=device),
                                     SplineCenter(times, knots, device=device))
    truth = FittableShiftedMultipole(FittableMultipole(target, device=device),
                                     SplineCenter(times, knots, device=device))
    with torch.no_grad():
        F_target = truth(xyz, t_eval)[1]
    return model, xyz, F_target


def evolving_halo_snapshots(n_snap: int = 8):
    """Growing triaxial halo: Dehnen (gamma=1, a=20 kpc) with mass 0.8 -> 1.0
    x 1e12 Msol and axes 1 : 0.8 -> 0.9 : 0.6 -> 0.7 over t in [0, 1]; lmax 8,
    50 nodes on [0.1, 500] kpc (coarse quadrature).  Returns (coefs, times)."""
    times = np.linspace(0.0, 1.0, n_snap)
    out = []
    for i in range(n_snap):
        f = i / max(1, n_snap - 1)
        rho = dehnen_density(1e12 * (0.8 + 0.2 * f), 20.0, 1.0, 0.8 + 0.1 * f, 0.6 + 0.1 * f)
        out.append(multipole_from_density(rho, 8, 0.1, 500.0, 50, symmetry="triaxial",
                                          n_theta=18, n_phi=26, fine_factor=6,
                                          r_in=1e-3, r_out=5e4))
    return out, times


def satellite_track(n_knots: int = 50, t_end: float = 1.2):
    """(T, 4) rows [t, x, y, z]: an infalling satellite, r = 100 - 40 t kpc,
    angle 2 pi t / 0.8 in a plane inclined by 0.3 rad."""
    t = np.linspace(0.0, t_end, n_knots)
    r = 100.0 - 40.0 * t
    w = 2.0 * math.pi / 0.8
    c, s = math.cos(0.3), math.sin(0.3)
    return np.stack([t, r * np.cos(w * t), r * np.sin(w * t) * c, r * np.sin(w * t) * s], axis=1)


def satellite_coefs():
    """Dehnen satellite (gamma=1, M=1e11, a=5 kpc, axes 1:0.9:0.8), lmax 4,
    30 nodes on [0.05, 200] kpc."""
    rho = dehnen_density(1e11, 5.0, 1.0, 0.9, 0.8)
    return multipole_from_density(rho, 4, 0.05, 200.0, 30, symmetry="triaxial",
                                  n_theta=14, n_phi=18, fine_factor=6, r_in=1e-4, r_out=2e4)


def leapfrog_problem(n: int, device=None, n_snap: int = 8, seed: int = 4):
    """Potential (Evolving halo of n_snap snapshots + Shifted lmax-4 satellite)
    and initial (x, v): Hernquist positions (a=20 kpc), isotropic Gaussian
    velocities with 120 km/s per component."""
    from . import EvolvingPotentialGPU, MultipolePotentialGPU, ShiftedPotentialGPU
    snaps, times = evolving_halo_snapshots(n_snap)
    halo = EvolvingPotentialGPU([MultipolePotentialGPU(c, device=device) for c in snaps], times)
    sat = ShiftedPotentialGPU(MultipolePotentialGPU(satellite_coefs(), device=device),
                              satellite_track())
    pot = halo + sat
    x = hernquist_points(n, seed)
    rng = np.random.default_rng(seed + 1)
    v = rng.normal(0.0, 120.0, (n, 3))
    return pot, x, v


def high_order_coefs(lmax: int = 32, n_r: int = 80, seed: int = 1) -> MultipoleCoefs:
    """High-order Multipole: a Hernquist density (M=1e12, a=20 kpc) with a
    seeded random triaxial stretch (axis ratios in [0.7, 1]) and a random
    angular perturbation sum_{l>=1,m} eps_lm 0.3^l Y_lm (l=1 included:
    lopsided), eps ~ 0.3 N(0,1); expanded to lmax on n_r log nodes over
    [0.1, 500] kpc (no symmetry: all (lmax+1)^2 harmonics)."""
    from .expansion import real_basis
    rng = np.random.default_rng(seed)
    p, q = np.sort(rng.uniform(0.7, 1.0, 2))[::-1]
    labels = [(l, m) for l in range(lmax + 1) for m in range(-l, l + 1)]
    eps = np.array([0.0 if l == 0 else 0.3 * rng.normal() * 0.3 ** l for l, m in labels])
    M, a = 1e12, 20.0

    cache = {}

    def angular(x, y, z):
        # the perturbation depends on the direction only; the quadrature
        # calls rho on (radii x directions) blocks sharing their directions
        key = x.shape[-1]
        if key not in cache:
            x0, y0, z0 = x.reshape(-1, key)[0], y.reshape(-1, key)[0], z.reshape(-1, key)[0]
            r0 = np.sqrt(x0 * x0 + y0 * y0 + z0 * z0)
            B = real_basis(lmax, z0 / r0, np.arctan2(y0, x0))
            cache[key] = 1.0 + eps @ B
        return cache[key]

    def rho(x, y, z):
        mm = np.sqrt(x * x + (y / p) ** 2 + (z / q) ** 2)
        base = M * a / (2.0 * math.pi * p * q) / (np.maximum(mm, 1e-300) * (mm + a) ** 3)
        return base * angular(x, y, z)

    return multipole_from_density(rho, lmax, 0.1, 500.0, n_r, symmetry="none",
                                  fine_factor=8, r_in=1e-3, r_out=5e4)


def mw_disk_bar_coefs(nR: int = 32, nz: int = 32, n_phi: int = 16):
    """CylSpline table of a Miyamoto-Nagai disk (M=6e10, a=3, b=0.3 kpc) plus
    a quadrupole bar (Dehnen 2000 form: R_b = 3.5 kpc, strength 0.01 of
    (220 km/s)^2 at R_0 = 8 kpc, angle 0.3 rad) with weaker m=4 and m=6
    companions, sampled on a 32 x 32 grid uniform in asinh(R/3), asinh(z/0.3)
    up to R = 200, z = 200 kpc (z >= 0 half; mirrored by the builder), and
    split into azimuthal harmonics m = 0..6 by an exact 16-point FFT."""
    from .coefs import CylSplineCoefs
    from .multipole_prep import G_AGAMA
    R = 3.0 * np.sinh(np.linspace(0.0, np.arcsinh(200.0 / 3.0), nR))
    z = 0.3 * np.sinh(np.linspace(0.0, np.arcsinh(200.0 / 0.3), nz))
    ph = np.arange(n_phi) * (2.0 * math.pi / n_phi)
    RR, ZZ, PP = np.meshgrid(R, z, ph, indexing="ij")
    Mg = G_AGAMA * 6e10
    phi_mn = -Mg / np.sqrt(RR * RR + (3.0 + np.sqrt(ZZ * ZZ + 0.09)) ** 2)
    r = np.sqrt(RR * RR + ZZ * ZZ)
    Rb, amp = 3.5, 0.01 * 220.0 ** 2 / 3.0 * (8.0 / 3.5) ** 3
    u = np.where(r >= Rb, -(np.maximum(r, Rb) / Rb) ** -3, (r / Rb) ** 3 - 2.0)
    ang = np.where(r > 0, (RR / np.maximum(r, 1e-300)) ** 2, 0.0)
    shape = amp * u * ang
    phi_bar = shape * (np.cos(2 * (PP - 0.3)) + 0.15 * np.cos(4 * (PP - 0.3)) * ang +
                       0.05 * np.cos(6 * (PP - 0.5)) * ang * ang)
    total = phi_mn + phi_bar
    F = np.fft.rfft(total, axis=2) / n_phi
    phi = {0: F[:, :, 0].real.copy()}
    floor = 1e-12 * np.max(np.abs(phi[0]))      # FFT round-off of absent harmonics
    for m in range(1, 7):
        c = 2.0 * F[:, :, m].real
        s = -2.0 * F[:, :, m].imag
        if np.max(np.abs(c)) > floor:
            phi[m] = c.copy()
        if np.max(np.abs(s)) > floor:
            phi[-m] = s.copy()
    return CylSplineCoefs(R, z, phi, 6)


def mw_composite(device=None):
    """Milky-Way-like composite: Multipole halo (halo_lmax8_coefs), CylSpline
    disk + bar (mw_disk_bar_coefs), NFW (M=1e12, rs=20 kpc), Plummer bulge
    (M=1e10, b=0.5 kpc)."""
    from . import (CylSplinePotentialGPU, MultipolePotentialGPU, NFWPotentialGPU,
                   PlummerPotentialGPU)
    return (MultipolePotentialGPU(halo_lmax8_coefs(), device=device) +
            CylSplinePotentialGPU(mw_disk_bar_coefs(), device=device) +
            NFWPotentialGPU(mass=1e12, scaleRadius=20.0, device=device) +
            PlummerPotentialGPU(mass=1e10, scaleRadius=0.5, device=device))
