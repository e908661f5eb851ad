"""ORACLE -- test infrastructure only (numpy restatement of the reference's
analytic potentials, _analytic_potentials.py:176-1165).

Each function returns (phi, grad, hess, rho) for (N,3) positions, where grad
is +grad Phi and hess the Hessian [xx, yy, zz, xy, yz, xz] -- the reference
kernels' raw outputs.  The formulas follow the cited CuPy kernel bodies; the
DiskAnsatz isothermal (h < 0) and thin (h = 0) vertical profiles follow
Agama's definitions (the reference uses the exponential profile for every h).
Also: the Composite/Shifted/Scaled/Evolving transforms (gpu_potential.py:
1439-1918) as plain sums over components.
"""
from __future__ import annotations

import math

import numpy as np

G = 4.30091727067736e-06
INV_4PIG = 1.0 / (4.0 * math.pi * G)
SLOTS = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]


def _radial(x, r, P, d1, d2):
    """Spherical Phi(r) -> grad, Hessian (reference NFW/Hernquist hess bodies)."""
    d1r = d1 / r
    A = (d2 - d1r) / (r * r)
    g = d1r[:, None] * x
    h = np.stack([A * x[:, i] * x[:, j] + (d1r if i == j else 0.0) for i, j in SLOTS], axis=1)
    return P, g, h


def _nfw_small(u):
    """Phi'(r) rs^2/GM and Phi''(r) rs^3/GM of the NFW potential for u = r/rs
    below 1/16, from the Taylor series of ln(1+u)/u (terms to u^16): the
    closed forms lose digits as 1/u and 1/u^2 there by cancellation."""
    d1 = sum((-1.0) ** n * (n - 1) / n * u ** (n - 2) for n in range(2, 19))
    d2 = sum((-1.0) ** n * (n - 1) * (n - 2) / n * u ** (n - 3) for n in range(3, 20))
    return d1, d2


def nfw(x, M, rs):                                     # :176-243
    GM = G * M
    r = np.linalg.norm(x, axis=1)
    r = np.where(r < 1e-300, 1e-300, r)                # reference clamp; NaN passes
    L = np.log1p(r / rs)
    rr = r + rs
    with np.errstate(over="ignore", invalid="ignore"):
        d1 = GM * (L / r ** 2 - 1.0 / (r * rr))
        d2 = GM * (-2.0 * L / r ** 3 + 1.0 / (r ** 2 * rr) + (2.0 * r + rs) / (r ** 2 * rr ** 2))
    small = r / rs < 0.0625
    if np.any(small):
        s1, s2 = _nfw_small(r[small] / rs)
        d1[small] = GM / rs ** 2 * s1
        d2[small] = GM / rs ** 3 * s2
    P, g, h = _radial(x, r, -GM * L / r, d1, d2)
    return P, g, h, M / (4.0 * math.pi * r * rr * rr)


def plummer(x, M, b):                                  # :288-340
    GM, b2 = G * M, b * b
    s2 = (x * x).sum(1) + b2
    s = np.sqrt(s2)
    c = GM / (s2 * s)
    g = c[:, None] * x
    A, B = -3.0 * GM / (s2 * s2 * s), GM / (s2 * s)
    h = np.stack([A * x[:, i] * x[:, j] + (B if i == j else 0.0) for i, j in SLOTS], axis=1)
    return -GM / s, g, h, 3.0 * M / (4.0 * math.pi) * b2 / s2 ** 2.5


def hernquist(x, M, a):                                # :383-444
    GM = G * M
    r = np.maximum(np.linalg.norm(x, axis=1), 1e-300)
    ra = r + a
    P, g, h = _radial(x, r, -GM / ra, GM / ra ** 2, -2.0 * GM / ra ** 3)
    return P, g, h, M * a / (2.0 * math.pi * r * ra ** 3)


def dehnen(x, M, a, gamma):                            # :486-549 (+ gamma = 2)
    GM = G * M
    r = np.maximum(np.linalg.norm(x, axis=1), 1e-300)
    ra = r + a
    e = 2.0 - gamma
    P = -(GM / a) * np.log1p(a / r) if e == 0 else -(GM / a) * (1.0 - (r / ra) ** e) / e
    d1 = GM * r ** (1.0 - gamma) * ra ** (gamma - 3.0)
    d2 = d1 * ((1.0 - gamma) / r + (gamma - 3.0) / ra)
    P, g, h = _radial(x, r, P, d1, d2)
    rho = (3.0 - gamma) * M / (4.0 * math.pi * a ** 3) * (r / a) ** (-gamma) * (1.0 + r / a) ** (gamma - 4.0)
    return P, g, h, rho


def isochrone(x, M, b):                                # :603-676
    GM = G * M
    r2 = (x * x).sum(1)
    s = np.sqrt(r2 + b * b)
    bs = b + s
    d1r = GM / (s * bs * bs)
    A = -GM * (b + 3.0 * s) / (s ** 3 * bs ** 3)
    g = d1r[:, None] * x
    h = np.stack([A * x[:, i] * x[:, j] + (d1r if i == j else 0.0) for i, j in SLOTS], axis=1)
    return -GM / bs, g, h, (A * r2 + 3.0 * d1r) * INV_4PIG


def miyamoto_nagai(x, M, a, b):                        # :720-793
    GM = G * M
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    R2 = X * X + Y * Y
    D = np.sqrt(Z * Z + b * b)
    aD = a + D
    S2 = R2 + aD * aD
    S = np.sqrt(S2)
    S3, S5 = S2 * S, S2 * S2 * S
    g = np.stack([GM * X / S3, GM * Y / S3, GM * aD * Z / (D * S3)], axis=1)
    dadz = Z / D
    term1 = ((dadz * D - aD * Z / D) * Z + aD * D) / (D * D * S3)
    term2 = 3.0 * aD * Z * (aD * Z / D) / (D * S5)
    off_z = -3.0 * GM * aD * Z / (D * S5)
    h = np.stack([GM * (1.0 / S3 - 3.0 * X * X / S5), GM * (1.0 / S3 - 3.0 * Y * Y / S5),
                  GM * (term1 - term2), -3.0 * GM * X * Y / S5, off_z * Y, off_z * X], axis=1)
    rho = GM * b * b * INV_4PIG * (a * R2 + (a + 3.0 * D) * aD * aD) / (D ** 3 * S2 ** 2.5)
    return -GM / S, g, h, rho


def loghalo(x, v0, rc, p, q):                          # :842-892
    v02, rc2, p2, q2 = v0 * v0, rc * rc, p * p, q * q
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    m2 = rc2 + X * X + Y * Y / p2 + Z * Z / q2
    m4 = m2 * m2
    c = v02 / m2
    g = np.stack([c * X, c * Y / p2, c * Z / q2], axis=1)
    k = -2.0 * v02 / m4
    h = np.stack([v02 * (1.0 / m2 - 2.0 * X * X / m4), v02 * (1.0 / (m2 * p2) - 2.0 * Y * Y / (m4 * p2 * p2)),
                  v02 * (1.0 / (m2 * q2) - 2.0 * Z * Z / (m4 * q2 * q2)), k * X * Y / p2,
                  k * Y * Z / (p2 * q2), k * X * Z / q2], axis=1)
    lap = v02 * ((1.0 + 1.0 / p2 + 1.0 / q2) / m2 - 2.0 * (X * X + Y * Y / p2 ** 2 + Z * Z / q2 ** 2) / m4)
    return 0.5 * v02 * np.log(m2), g, h, lap * INV_4PIG


def _exp_m1_plus_series(u):
    """e^-u - 1 + u = sum_{k>=2} (-u)^k / k! summed in long double to k = 25."""
    u = np.asarray(u, np.longdouble)
    out = np.zeros_like(u)
    term = u * u / 2
    for k in range(2, 26):
        out += term
        term = term * -u / (k + 1)
    return out.astype(np.float64)


def disk_ansatz(x, Sigma, Rd, hz, rin=0.0):            # :950-1052
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    r = np.maximum(np.linalg.norm(x, axis=1), 1e-300)
    ir = 1.0 / r
    f = 4.0 * math.pi * G * Sigma * np.exp(-r / Rd - rin / r)
    fo = -1.0 / Rd + rin * ir * ir
    f1, f2 = f * fo, f * (fo * fo - 2.0 * rin * ir ** 3)
    az, sg = np.abs(Z), np.sign(Z)
    if hz > 0:
        # e^-u - 1 + u and 1 - e^-u without the cancellation near the midplane
        # (the reference's form keeps 1e-16 / u^2): a Taylor sum for u < 0.5
        u = az / hz
        g = np.where(u >= 0.5, u + np.expm1(-u), _exp_m1_plus_series(u))
        Hv, H1, H2 = 0.5 * hz * g, -0.5 * sg * np.expm1(-u), 0.5 * np.exp(-u) / hz
    elif hz < 0:                                      # Agama isothermal: |h|/2 ln cosh(z/|h|)
        h = -hz
        u = az / h
        us = np.minimum(u, 1.0)                       # (the branch np.where discards)
        lc = np.where(u < 1.0, np.log1p(2.0 * np.sinh(0.5 * us) ** 2),
                      u + np.log1p(np.exp(-2.0 * np.maximum(u, 1.0))) - math.log(2.0))
        Hv = 0.5 * h * lc                             # ln cosh without overflow or cancellation
        th = np.tanh(Z / h)
        H1, H2 = 0.5 * th, 0.5 * (1.0 - th * th) / h
    else:                                             # Agama thin disk: |z|/2
        Hv, H1, H2 = 0.5 * az, 0.5 * sg, 0.0 * az
    common = Hv * f1 * ir
    g = np.stack([common * X, common * Y, common * Z + H1 * f], axis=1)
    xr, yr, zr = X * ir, Y * ir, Z * ir
    A, B = Hv * (f2 - f1 * ir), Hv * f1 * ir
    h = np.stack([A * xr * xr + B, A * yr * yr + B,
                  Hv * (f2 * zr * zr + f1 * (1.0 - zr * zr) * ir) + 2.0 * f1 * H1 * zr + f * H2,
                  A * xr * yr, A * yr * zr + f1 * H1 * yr, A * xr * zr + f1 * H1 * xr], axis=1)
    lap = Hv * (f2 + 2.0 * f1 * ir) + 2.0 * f1 * H1 * zr + f * H2
    return f * Hv, g, h, lap * INV_4PIG


def uniform_accel(x, ax, ay, az):                      # :1112-1165
    n = len(x)
    P = -(ax * x[:, 0] + ay * x[:, 1] + az * x[:, 2])
    g = np.tile(np.array([-ax, -ay, -az]), (n, 1))
    return P, g, np.zeros((n, 6)), np.zeros(n)


TYPES = {
    "NFW": (nfw, ("mass", "scaleRadius")),
    "Plummer": (plummer, ("mass", "scaleRadius")),
    "Hernquist": (hernquist, ("mass", "scaleRadius")),
    "Dehnen": (dehnen, ("mass", "scaleRadius", "gamma")),
    "Isochrone": (isochrone, ("mass", "scaleRadius")),
    "MiyamotoNagai": (miyamoto_nagai, ("mass", "scaleRadius", "scaleHeight")),
    "LogHalo": (loghalo, ("velocity", "coreRadius", "axisRatioY", "axisRatioZ")),
    "DiskAnsatz": (disk_ansatz, ("surfaceDensity", "scaleRadius", "scaleHeight", "innerCutoffRadius")),
    "UniformAcceleration": (uniform_accel, ("ax", "ay", "az")),
}


def evaluate(type_name: str, x, **kw):
    fn, names = TYPES[type_name]
    return fn(np.asarray(x, np.float64), *[kw[n] for n in names if n in kw])


def transformed(fn, x, ampl=1.0, sigma=1.0, tau=(0.0, 0.0, 0.0)):
    """Phi = ampl * f(sigma (x - tau)) and its derivatives by the chain rule
    (Shifted gpu_potential.py:1738-1764, Scaled 1878-1900)."""
    P, g, h, rho = fn(sigma * (np.asarray(x) - np.asarray(tau)))
    return ampl * P, ampl * sigma * g, ampl * sigma ** 2 * h, ampl * sigma ** 2 * rho
