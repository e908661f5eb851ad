"""Pins the CylSpline oracle on random table structures (tests/random_tables.py:
mmax 0..6 with even, cos-only, all or random subsets of harmonics, half grids
mirrored by the build and full z-asymmetric grids).

* node exactness: at an interior grid node (R_i, z_j) and any azimuth the
  oracle's Phi is sum_m Phi_m(R_i, z_j) T_m(phi) with the table's own values
  (the bicubic spline interpolates its nodes; log-scaling is undone exactly);
* calculus: the oracle's gradient is the central difference of its Phi, inside
  the grid box and outside it (the exterior multipole).
"""
import numpy as np
import pytest

from oracle.cylspline_oracle import OracleCylSpline
from random_tables import cyl_sweep_points, random_cylspline

SEEDS = list(range(12))


def _table_value(c, i, j, ph):
    """sum_m Phi_m(R_i, z_j) T_m(phi) over the table's harmonics; for a half
    grid row j counts from z = 0 (mirror images are the same values)."""
    out = np.zeros_like(ph)
    for m, t in c.phi.items():
        T = np.cos(m * ph) if m >= 0 else np.sin(-m * ph)
        out += t[i, j] * T
    return out


@pytest.mark.parametrize("seed", SEEDS)
def test_random_cylspline_node_exactness(seed):
    c, d = random_cylspline(seed)
    o = OracleCylSpline(c)
    ph = np.random.default_rng(seed).uniform(0, 2 * np.pi, 7)
    for i in sorted({1, d["nR"] // 2, d["nR"] - 2}):
        for j in sorted({1, len(c.z_grid) // 2, len(c.z_grid) - 2}):
            R, z = c.R_grid[i], c.z_grid[j]
            x = np.stack([R * np.cos(ph), R * np.sin(ph), np.full_like(ph, z)], axis=1)
            P = o.eval(x)[0]
            ex = _table_value(c, i, j, ph)
            assert np.max(np.abs(P - ex)) <= 1e-12 * np.max(np.abs(ex)), (d, i, j)


@pytest.mark.parametrize("seed", SEEDS[::2])
def test_random_cylspline_gradient(seed):
    c, d = random_cylspline(seed)
    o = OracleCylSpline(c)
    x0 = cyl_sweep_points(d, 400, seed + 11)[100:]      # off the axis
    P0, g0, _, outer = o.eval(x0)
    h = 1e-6 * np.maximum(np.linalg.norm(x0, axis=1), 1e-3)
    gd = np.zeros_like(g0)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1.0
        gd[:, k] = (o.eval(x0 + h[:, None] * e)[0] - o.eval(x0 - h[:, None] * e)[0]) / (2 * h)
    for part in (outer, ~outer):
        if part.any():
            scale = np.abs(g0[part]).max()
            assert np.max(np.abs(g0[part] - gd[part])) <= 1e-5 * scale, d
