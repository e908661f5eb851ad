"""Potentials whose values the mathematics fixes, for pinning the
extrapolation branches of the Multipole and CylSpline paths.

Multipole monopole tables sampled at log-spaced nodes from closed forms that
the l=0 extrapolation laws reproduce exactly:

* ``inner_log``   Phi = W + U ln(r / r_min): the inner law (v = 0) in its
  logarithmic (s == v) form, the form the reference's Lambert-W branch choice
  yields (gpu_potential.py:217-242);
* ``inner_core``  Phi = W + U (r / r_min)^2: the s = 2 constant-density core,
  selected by the cubic-versus-power-law test of dPhi/dr at the second node
  (gpu_potential.py:245-251);
* ``outer_log``   Phi = (W + U ln(r / r_max)) r_max / r: the outer law
  (v = -1) in its s == v form (gpu_potential.py:313-324);
* ``kepler``      Phi = -GM / r: the round-off fallback to a pure Keplerian
  tail (gpu_potential.py:297-306);
* ``outer_power`` Phi = W r_max / r + U (r / r_max)^-2.5: not reproduced --
  the reference's branch choice gives s = v, so beyond r_max the potential is
  the log law through Phi(r_max) and r dPhi/dr(r_max):
  Phi = (Phi_N + (r_N Phi'_N + Phi_N) ln(r / r_N)) r_N / r.

Each entry gives Phi, dPhi/dr and d2Phi/dr2 in closed form and the region
(``"inner"``: r < r_min, ``"outer"``: r > r_max) where they must hold.

CylSpline: an exterior field written as Cartesian solid harmonics (monopole,
dipole, quadrupole m = 0 and m = +-2, octupole m = 0), whose PowerLaw
coefficients W_lm follow from Agama's normalisation by hand (below), so a fit
on the grid boundary must recover them.
"""
from __future__ import annotations

import math

import numpy as np


def _grid(rmin, rmax, n):
    return np.exp(np.linspace(math.log(rmin), math.log(rmax), n))


def radial_law_cases():
    """name -> (r_grid, Phi(r), dPhi/dr(r), d2Phi/dr2(r), region)."""
    out = {}
    W, U, r0 = -100.0, 3.0, 0.1
    out["inner_log"] = (_grid(r0, 10.0, 40),
                        lambda r: W + U * np.log(r / r0), lambda r: U / r,
                        lambda r: -U / r ** 2, "inner")
    Uc = 0.005
    out["inner_core"] = (_grid(r0, 10.0, 40),
                         lambda r: W + Uc * (r / r0) ** 2, lambda r: 2 * Uc * r / r0 ** 2,
                         lambda r: 2 * Uc / r0 ** 2 + 0 * r, "inner")
    Wo, Uo, rN = -50.0, 8.0, 100.0
    out["outer_log"] = (_grid(1.0, rN, 40),
                        lambda r: (Wo + Uo * np.log(r / rN)) * rN / r,
                        lambda r: (Uo - Wo - Uo * np.log(r / rN)) * rN / r ** 2,
                        lambda r: (3 * Uo - 2 * Wo - 2 * Uo * np.log(r / rN)) * rN / r ** 3 * -1.0
                        + 0 * r, "outer")
    GM = 1000.0
    out["kepler"] = (_grid(0.1, rN, 40), lambda r: -GM / r, lambda r: GM / r ** 2,
                     lambda r: -2 * GM / r ** 3, "outer")
    Wp, Up, sp = -50.0, -5.0, -2.5
    P_N = Wp + Up
    D_N = (-Wp + sp * Up)            # r dPhi/dr at r_N
    Ul, Wl = D_N + P_N, P_N
    out["outer_power"] = (_grid(1.0, rN, 40),
                          lambda r: (Wl + Ul * np.log(r / rN)) * rN / r,
                          lambda r: (Ul - Wl - Ul * np.log(r / rN)) * rN / r ** 2,
                          lambda r: -(3 * Ul - 2 * Wl - 2 * Ul * np.log(r / rN)) * rN / r ** 3,
                          "outer")
    # the table itself samples the power law (the fixture's nodes), not the log law
    out["outer_power"] = out["outer_power"] + (
        lambda r: Wp * rN / r + Up * (r / rN) ** sp,
        lambda r: -Wp * rN / r ** 2 + sp * Up * (r / rN) ** sp / r)
    return out


def radial_table(name):
    """MultipoleCoefs (monopole) sampled at the case's nodes."""
    from agama_gpu_amd.coefs import MultipoleCoefs
    case = radial_law_cases()[name]
    r = case[0]
    f, df = (case[5], case[6]) if len(case) > 5 else (case[1], case[2])
    return MultipoleCoefs(r, f(r)[:, None], df(r)[:, None], [(0, 0)], 0)


def radial_points(name, n, seed):
    """Points in the case's extrapolated region, isotropic directions."""
    case = radial_law_cases()[name]
    r_grid, region = case[0], case[4]
    rng = np.random.default_rng(seed)
    if region == "inner":
        r = r_grid[0] * np.exp(rng.uniform(math.log(1e-4), math.log(0.999999), n))
    else:
        r = r_grid[-1] * np.exp(rng.uniform(math.log(1.000001), math.log(1e3), n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    return u * r[:, None]


def radial_exact(name, x):
    """Closed-form (Phi, force, deriv) of the case's extrapolation law at x
    (deriv in the slots [xx, yy, zz, xy, yz, xz], force = -grad, deriv = -H)."""
    case = radial_law_cases()[name]
    r = np.linalg.norm(x, axis=1)
    P, d1, d2 = case[1](r), case[2](r), case[3](r)
    u = x / r[:, None]
    F = -d1[:, None] * u
    slots = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    H = np.stack([(d2 - d1 / r) * u[:, i] * u[:, j] + (d1 / r if i == j else 0.0)
                  for i, j in slots], axis=1)
    return P, F, -H


# ---- CylSpline: exterior field of solid harmonics -----------------------------
# Phi = -GM/r + D x/r^3 + Q (3z^2 - r^2)/r^5 + Q2 (x^2 - y^2)/r^5
#       + Q3 x y / r^5 + Q4 z (5 z^2 - 3 r^2) / r^7
# Agama's PowerLaw term  W_lm (r0/r)^(l+1) mul_m Ptilde_l^|m|(cos th) T_m(phi),
# mul_0 = 2 sqrt(pi), mul_m = 2 sqrt(2 pi), Ptilde = sqrt((2l+1)/(4 pi)
# (l-|m|)!/(l+|m|)!) P_l^|m| with the Condon-Shortley phase, gives
#   mul_0 Ptilde_00 = 1                      -> W_00  = -GM / r0
#   mul_1 Ptilde_11 = -sqrt(3) sin th        -> W_11  = -D / (sqrt(3) r0^2)
#   mul_0 Ptilde_20 = sqrt(5) (3c^2 - 1) / 2 -> W_20  = 2 Q / (sqrt(5) r0^3)
#   mul_2 Ptilde_22 = sqrt(15) / 2 sin^2 th  -> W_22  = 2 Q2 / (sqrt(15) r0^3),
#                 (x y = R^2 sin 2phi / 2)   -> W_2-2 = Q3 / (sqrt(15) r0^3)
#   mul_0 Ptilde_30 = sqrt(7) P_3(c)         -> W_30  = 2 Q4 / (sqrt(7) r0^4)
EXT = dict(GM=1.0, D=0.02, Q=0.05, Q2=0.03, Q3=0.02, Q4=0.01)


def exterior_phi(x, y, z, p=EXT):
    r2 = x * x + y * y + z * z
    r = np.sqrt(r2)
    return (-p["GM"] / r + p["D"] * x / r ** 3 + p["Q"] * (3 * z * z - r2) / r ** 5
            + p["Q2"] * (x * x - y * y) / r ** 5 + p["Q3"] * x * y / r ** 5
            + p["Q4"] * z * (5 * z * z - 3 * r2) / r ** 7)


def exterior_harmonics(R, z, p=EXT):
    """Phi_m(R, z) of the field for m = -2..2 (T_m = cos m phi, sin |m| phi)."""
    r2 = R * R + z * z
    r = np.sqrt(r2)
    return {0: -p["GM"] / r + p["Q"] * (3 * z * z - r2) / r ** 5
               + p["Q4"] * z * (5 * z * z - 3 * r2) / r ** 7,
            1: p["D"] * R / r ** 3, -1: 0.0 * r,
            2: p["Q2"] * R * R / r ** 5, -2: p["Q3"] * R * R / (2 * r ** 5)}


def exterior_W(r0, p=EXT, lmax=8):
    W = np.zeros((lmax + 1) ** 2)
    W[0] = -p["GM"] / r0
    W[1 * 2 + 1] = -p["D"] / (math.sqrt(3) * r0 ** 2)
    W[2 * 3 + 0] = 2 * p["Q"] / (math.sqrt(5) * r0 ** 3)
    W[2 * 3 + 2] = 2 * p["Q2"] / (math.sqrt(15) * r0 ** 3)
    W[2 * 3 - 2] = p["Q3"] / (math.sqrt(15) * r0 ** 3)
    W[3 * 4 + 0] = 2 * p["Q4"] / (math.sqrt(7) * r0 ** 4)
    return W


def exterior_grid():
    """R in [0, 60] (20 nodes), z in [-80, 80] (26 nodes, no z = 0 node)."""
    R = np.concatenate([[0.0], 60.0 * np.sinh(np.linspace(0.05, 2.0, 19)) / math.sinh(2.0)])
    R[-1] = 60.0
    zp = 80.0 * np.sinh(np.linspace(0.02, 2.0, 13)) / math.sinh(2.0)
    zp[-1] = 80.0
    return R, np.concatenate([-zp[::-1], zp])


def exterior_coefs(p=EXT):
    """CylSplineCoefs of the field on exterior_grid.  Nodes at r < 2 take the
    field at r = 2 along the same direction (a finite stand-in for the source
    region; the fit uses boundary nodes only), so every m = 0 node is
    negative and the builder log-scales."""
    from agama_gpu_amd.coefs import CylSplineCoefs
    R, z = exterior_grid()
    RR, ZZ = np.meshgrid(R, z, indexing="ij")
    s = np.maximum(1.0, 2.0 / np.hypot(RR, ZZ))
    tabs = exterior_harmonics(RR * s, ZZ * s, p)
    return CylSplineCoefs(R, z, {m: np.asarray(t, float) for m, t in tabs.items()}, 2)


def axis_tables():
    """Exterior fields with m = 0 and m = +-2 harmonics only (no dipole), so
    the spline's potential is twice differentiable across the z axis:
    log-scaled (every m = 0 node negative) and, shifted by a constant,
    linear."""
    from agama_gpu_amd.coefs import CylSplineCoefs
    c = exterior_coefs(dict(EXT, D=0.0))
    lin = CylSplineCoefs(c.R_grid, c.z_grid,
                         {m: t + (0.3 if m == 0 else 0.0) for m, t in c.phi.items()}, c.mmax)
    return {"log": c, "linear": lin}
