"""Pins the Multipole oracle on random table structures (tests/random_tables.py).

The fixtures cover three tables; these cover lmax 0..12 with sin terms, odd l
and m, sparse random subsets, even-even (stride-2) programs and tables whose
monopole changes sign (no log-scaling, constant inner law).  For each:

* node exactness: at r = r_i the oracle's Phi and dPhi/dr equal the file's
  columns times B_lm(angles) built independently from scipy's sph_harm_y;
* calculus: the oracle's force is minus the central difference of its Phi,
  its Hessian minus the central difference of its force, at random points
  inside the grid and in both extrapolation regions.
"""
import numpy as np
import pytest

from oracle.oracle import OracleMultipole
from random_tables import random_multipole
from test_oracle_multipole import basis, unit_vectors

SEEDS = list(range(18))


@pytest.mark.parametrize("seed", SEEDS)
def test_random_table_node_exactness(seed):
    c, d = random_multipole(seed)
    o = OracleMultipole.from_coefs(c)
    u = unit_vectors(30, seed)
    B = basis(c.lm_labels, u)
    # interior nodes: r_max itself (and r_min where rounding puts |x| below it)
    # takes the l = 0 extrapolation law, as in the reference
    for i in sorted({1, d["nR"] // 2, d["nR"] - 2}):
        r = c.R_grid[i]
        p, F = o.pot_force(u * r)
        ex_p = B @ c.phi[i]
        ex_dr = B @ c.dphi_dr[i]
        dr = -(F * u).sum(1)
        assert np.max(np.abs(p - ex_p)) / np.max(np.abs(ex_p)) < 1e-12, (d, i)
        assert np.max(np.abs(dr - ex_dr)) / np.max(np.abs(ex_dr)) < 1e-10, (d, i)


@pytest.mark.parametrize("seed", SEEDS[::3])
@pytest.mark.parametrize("region", ["inner", "grid", "outer"])
def test_random_table_calculus(seed, region):
    c, d = random_multipole(seed)
    o = OracleMultipole.from_coefs(c)
    rng = np.random.default_rng(seed)
    lo, hi = {"inner": (d["r_min"] / 10, d["r_min"] / 1.5),
              "grid": (d["r_min"] * 1.5, d["r_max"] / 1.5),
              "outer": (d["r_max"] * 1.5, d["r_max"] * 10)}[region]
    r = np.exp(rng.uniform(np.log(lo), np.log(hi), 12))
    x0 = unit_vectors(12, seed + 1) * r[:, None]
    # keep away from the z axis, where the finite differences lose digits
    x0[:, 0] += 0.3 * r
    p0, F0, D0 = o.eval_deriv(x0)
    h = 1e-5 * np.linalg.norm(x0, axis=1)
    Fd = np.zeros_like(F0)
    Hd = np.zeros((len(x0), 3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1.0
        xp, xm = x0 + h[:, None] * e, x0 - h[:, None] * e
        Fd[:, k] = -(o.potential(xp) - o.potential(xm)) / (2 * h)
        Hd[:, :, k] = -(o.pot_force(xp)[1] - o.pot_force(xm)[1]) / (2 * h[:, None])
    # (<=: a constant inner law has zero force and Hessian on both sides)
    assert np.max(np.abs(F0 - Fd)) <= 1e-6 * np.max(np.abs(F0)), d
    slots = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    Dd = np.stack([Hd[:, i, j] for i, j in slots], axis=1)
    # Hd is the Hessian of Phi (minus the difference of the force); deriv = -H
    assert np.max(np.abs(D0 + Dd)) <= 1e-5 * np.max(np.abs(D0)), d
