"""Pins the Multipole oracle (oracle/mp_oracle.c) without Agama.

Each check is chosen so a plausible mistake (dropped term, wrong sign or
index, transposed operand, wrong normalisation) fails it:

* node exactness: at r = r_i the spline reproduces the file's Phi_lm(r_i) and
  dPhi_lm/dr(r_i), so Phi and dPhi/dr equal sum_lm value * B_lm(angles) with
  B_lm built independently from scipy.special.sph_harm_y;
* the table builder agrees with the independent numpy/scipy restatement in
  the product (two codes, same reference algorithm);
* calculus: force = -grad Phi and deriv = -grad force by central differences,
  density = trace(Hessian)/(4 pi G);
* closed forms: a Hernquist monopole table reproduces -GM/(r+a) to spline
  accuracy; the quintic spline reproduces a quintic polynomial exactly;
* the l=0 extrapolation laws are continuous (value and slope) at r_min, r_max;
* rotation about z by 90 degrees maps (l, m) terms by known signs.
"""
import math

import numpy as np
import pytest
from scipy.special import sph_harm_y

from conftest import FIXTURES
from oracle.oracle import INV_4PIG, OracleMultipole
from agama_gpu_amd.coefs import MultipoleCoefs, read_coefs
from agama_gpu_amd.multipole_prep import build_multipole_data

G = 4.30091727067736e-06
FILES = ["600.dark.none_8.coef_mul_DR", "100.LMC.none_8.coef_mult"]


def basis(labels, u):
    th = np.arccos(u[:, 2])
    ph = np.arctan2(u[:, 1], u[:, 0])
    cols = []
    for l, m in labels:
        am = abs(m)
        Pt = np.real(sph_harm_y(l, am, th, ph) * np.exp(-1j * am * ph))
        mul = 2 * np.sqrt(np.pi) if m == 0 else 2 * np.sqrt(2 * np.pi)
        T = np.cos(am * ph) if m >= 0 else np.sin(am * ph)
        cols.append(mul * Pt * T)
    return np.asarray(cols).T


def unit_vectors(n, seed):
    u = np.random.default_rng(seed).normal(size=(n, 3))
    return u / np.linalg.norm(u, axis=1)[:, None]


@pytest.mark.parametrize("fname", FILES)
def test_node_exactness(fname):
    c = read_coefs(FIXTURES / fname)
    o = OracleMultipole.from_coefs(c)
    u = unit_vectors(40, 3)
    B = basis(c.lm_labels, u)
    for i in (1, 7, 12, 20, 23):
        r = c.R_grid[i]
        p, F = o.pot_force(u * r)
        ex_p = B @ c.phi[i]
        ex_dr = B @ c.dphi_dr[i]
        dr = -(F * u).sum(1)                    # radial derivative of Phi
        # the file's grid is log-uniform to ~6e-13 only, which bounds exactness
        assert np.max(np.abs(p - ex_p)) / np.max(np.abs(ex_p)) < 1e-10
        assert np.max(np.abs(dr - ex_dr)) / np.max(np.abs(ex_dr)) < 1e-8


@pytest.mark.parametrize("fname", FILES)
def test_table_builder_matches_product_restatement(fname):
    c = read_coefs(FIXTURES / fname)
    o = OracleMultipole.from_coefs(c)
    d = build_multipole_data(c)
    assert np.array_equal(o.lm_l, d["lm_l"]) and np.array_equal(o.lm_m, d["lm_m"])
    scale = np.abs(d["poly"]).max()
    assert np.abs(o.poly - d["poly"]).max() < 1e-13 * scale
    for k in ("invPhi0", "inner_s", "inner_U", "inner_W", "outer_s", "outer_U", "outer_W",
              "logr_min", "dlogr", "log_scaling", "inner_log", "outer_log"):
        assert o.scalars[k] == pytest.approx(d[k], rel=1e-13, abs=1e-300), k


def _fd(f, x, h):
    out = []
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        out.append((f(x + e) - f(x - e)) / (2 * h))
    return out


@pytest.mark.parametrize("fname", FILES)
def test_calculus_consistency(fname):
    c = read_coefs(FIXTURES / fname)
    o = OracleMultipole.from_coefs(c)
    rng = np.random.default_rng(5)
    u = unit_vectors(200, 7)
    r = np.exp(rng.uniform(np.log(0.3), np.log(300.0), 200))
    x = u * r[:, None]
    p, F, D = o.eval_deriv(x)
    h = 1e-5 * r[:, None]
    g = np.stack([(o.potential(x + h * e) - o.potential(x - h * e)) / (2 * h[:, 0])
                  for e in np.eye(3)], axis=1)
    assert np.max(np.abs(g + F)) / np.max(np.abs(F)) < 1e-6
    # deriv = d force / d x, slots [xx, yy, zz, xy, yz, xz]
    cols = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    Fp = [o.pot_force(x + h * e)[1] for e in np.eye(3)]
    Fm = [o.pot_force(x - h * e)[1] for e in np.eye(3)]
    for slot, (i, j) in enumerate(cols):
        fd = (Fp[j][:, i] - Fm[j][:, i]) / (2 * h[:, 0])
        assert np.max(np.abs(fd - D[:, slot])) / np.max(np.abs(D)) < 1e-5, slot
    rho = o.density(x)
    trace = -(D[:, 0] + D[:, 1] + D[:, 2]) * INV_4PIG
    assert np.max(np.abs(rho - trace)) / np.max(np.abs(trace)) < 1e-9


def test_hernquist_monopole_closed_form():
    """Monopole table from -GM/(r+a): interpolation error is small and falls
    at the quintic spline's order when the grid is refined."""
    M, a = 1e12, 20.0
    u = unit_vectors(500, 11)
    rr = np.exp(np.random.default_rng(2).uniform(np.log(0.06), np.log(700.0), 500))
    x = u * rr[:, None]
    ex = -G * M / (rr + a)
    exF = -(G * M / (rr + a) ** 2)[:, None] * u
    errs = []
    for n in (60, 120):
        r = np.exp(np.linspace(np.log(0.05), np.log(800.0), n))
        c = MultipoleCoefs(r, (-G * M / (r + a))[:, None], (G * M / (r + a) ** 2)[:, None],
                           [(0, 0)], 0)
        o = OracleMultipole.from_coefs(c)
        p, F = o.pot_force(x)
        errs.append((np.max(np.abs(p / ex - 1)),
                     np.max(np.abs(F - exF)) / np.max(np.abs(exF))))
    assert errs[1][0] < 1e-8 and errs[1][1] < 1e-6
    # halving h cuts the error by ~2^5 (value) / 2^4 (derivative) or more
    assert errs[0][0] / errs[1][0] > 16 and errs[0][1] / errs[1][1] > 8


def test_quintic_reproduces_cubic():
    """The C2 quintic spline with vanishing 4th derivative at the ends
    reproduces any cubic in ln r exactly (every interval carries the same
    cubic); a dropped or mis-signed term of the tridiagonal system breaks it."""
    r = np.exp(np.linspace(0.0, 3.0, 16))
    t = np.log(r)
    cf = [10.3, -1.2, 0.7, 0.25]
    f = sum(cf[j] * t ** j for j in range(4))
    df = sum(j * cf[j] * t ** (j - 1) for j in range(1, 4))
    c = MultipoleCoefs(r, f[:, None], (df / r)[:, None], [(0, 0)], 0)
    o = OracleMultipole.from_coefs(c)
    tt = np.linspace(0.01, 2.99, 300)
    x = np.stack([np.exp(tt), 0 * tt, 0 * tt], axis=1)
    ex = sum(cf[j] * tt ** j for j in range(4))
    exd = sum(j * cf[j] * tt ** (j - 1) for j in range(1, 4))
    p, F = o.pot_force(x)
    assert np.max(np.abs(p - ex)) < 1e-12
    assert np.max(np.abs(-F[:, 0] * np.exp(tt) - exd)) < 1e-11


@pytest.mark.parametrize("fname", FILES)
def test_extrapolation_continuity(fname):
    c = read_coefs(FIXTURES / fname)
    o = OracleMultipole.from_coefs(c)
    # l=0 only beyond the grid, so compare the angular average (the monopole)
    u = np.array([[1.0, 0, 0], [-1, 0, 0], [0, 1.0, 0], [0, -1, 0], [0, 0, 1.0], [0, 0, -1]])
    for r0, sgn in ((c.R_grid[0], 1), (c.R_grid[-1], -1)):
        for eps in (1e-9,):
            pin = o.potential(u * r0 * (1 + sgn * eps)).mean()
            pout = o.potential(u * r0 * (1 - sgn * eps)).mean()
            # a 6-point average removes l=1,2 odd parts; compare to the monopole scale
            assert abs(pin - pout) / abs(pout) < 2e-2
    # monopole-only table: continuity is exact up to the spline derivative
    r = c.R_grid
    c0 = MultipoleCoefs(r, c.phi[:, :1], c.dphi_dr[:, :1], [(0, 0)], 0)
    o0 = OracleMultipole.from_coefs(c0)
    for r0 in (r[0], r[-1]):
        a = o0.potential(np.array([[r0 * (1 - 1e-10), 0, 0]]))[0]
        b = o0.potential(np.array([[r0 * (1 + 1e-10), 0, 0]]))[0]
        assert abs(a - b) < 1e-8 * abs(a)
        fa = o0.pot_force(np.array([[r0 * (1 - 1e-10), 0, 0]]))[1][0, 0]
        fb = o0.pot_force(np.array([[r0 * (1 + 1e-10), 0, 0]]))[1][0, 0]
        assert abs(fa - fb) < 1e-6 * abs(fa)


def test_rotation_about_z():
    # rotating the point by +90 deg about z equals rotating the table by -90 deg:
    # cos(m(phi+pi/2)) = cos(m phi)cos(m pi/2) - sin(m phi) sin(m pi/2)
    c = read_coefs(FIXTURES / FILES[0])
    o = OracleMultipole.from_coefs(c)
    phi = c.phi.copy()
    dphi = c.dphi_dr.copy()
    idx = {lm: j for j, lm in enumerate(c.lm_labels)}
    for (l, m), j in idx.items():
        am = abs(m)
        if am == 0:
            continue
        cm, sm = round(math.cos(am * math.pi / 2)), round(math.sin(am * math.pi / 2))
        jc, js = idx[(l, am)], idx[(l, -am)]
        phi[:, jc] = c.phi[:, jc] * cm + c.phi[:, js] * sm
        phi[:, js] = -c.phi[:, jc] * sm + c.phi[:, js] * cm
        dphi[:, jc] = c.dphi_dr[:, jc] * cm + c.dphi_dr[:, js] * sm
        dphi[:, js] = -c.dphi_dr[:, jc] * sm + c.dphi_dr[:, js] * cm
    o2 = OracleMultipole(c.R_grid, phi, dphi, c.lm_labels)
    x = unit_vectors(100, 4) * np.exp(np.random.default_rng(9).uniform(0, 5, 100))[:, None]
    xr = np.stack([-x[:, 1], x[:, 0], x[:, 2]], axis=1)     # +90 deg
    assert np.max(np.abs(o.potential(xr) - o2.potential(x))) < 1e-9 * np.abs(o.potential(x)).max()


def test_origin_and_pole_branches():
    c = read_coefs(FIXTURES / FILES[0])
    o = OracleMultipole.from_coefs(c)
    d = build_multipole_data(c)
    p, F = o.pot_force(np.zeros((1, 3)))
    assert p[0] == pytest.approx(d["inner_W"], rel=1e-15) and np.all(F == 0)
    # approach to the z axis is continuous, every force component, at both
    # poles (reference tests/test_gpu_audit.py:176-203): the m = 1 terms' limit
    # of Y / sin(theta) is +dY/dtheta at the north pole and -dY/dtheta at the
    # south pole (the reference's +dY/dtheta at both reverses the m = 1 part of
    # F_x, F_y at (0, 0, z < 0))
    for z in (5.0, -5.0, 40.0, -40.0):
        ph = np.linspace(0.0, 2 * np.pi, 8, endpoint=False)
        near = np.stack([1e-7 * np.cos(ph), 1e-7 * np.sin(ph), z + 0 * ph], axis=1)
        on = np.array([[0.0, 0.0, z]])
        pn, Fn = o.pot_force(near)
        po, Fo = o.pot_force(on)
        assert np.max(np.abs(pn - po[0])) < 1e-9 * abs(po[0])
        assert np.max(np.abs(Fn - Fo[0])) < 1e-6 * np.abs(Fo).max(), z
        assert np.abs(Fo[0, :2]).max() > 1e-3 * abs(Fo[0, 2])     # m = 1 terms present


def _fd_hessian(o, x0, h):
    """+Hessian at x0 (1, 3) by a 5-point central difference of the force."""
    H = np.zeros((3, 3))
    for j in range(3):
        e = np.zeros((1, 3))
        e[0, j] = h
        F = [o.pot_force(x0 + k * e)[1][0] for k in (2, 1, -1, -2)]
        H[:, j] = -(-F[0] + 8 * F[1] - 8 * F[2] + F[3]) / (12 * h)
    return H


def _as_matrix(D):
    """deriv [xx, yy, zz, xy, yz, xz] (= -Hessian) -> +Hessian 3 x 3."""
    return -np.array([[D[0], D[3], D[5]], [D[3], D[1], D[4]], [D[5], D[4], D[2]]])


@pytest.mark.parametrize("fname", FILES + ["halo"])
def test_hessian_on_the_axis(fname):
    """On the z axis (and within sin(theta) < 1e-10 of it) the Hessian is the
    regular limit: it equals a 5-point difference of the force, which is
    evaluated off the axis by the ordinary spherical chain rule, at both
    poles; the reference's pole branch (d2Y/dtheta2 = 0, no |m| = 1, 2 limits)
    was 10-40 % of max|H| off here.  H_xx != H_yy and H_xy, H_xz, H_yz != 0
    where m = 1, 2 harmonics are present."""
    if fname == "halo":
        from agama_gpu_amd.workloads import halo_lmax8_coefs
        o = OracleMultipole.from_coefs(halo_lmax8_coefs())
    else:
        o = OracleMultipole.from_coefs(read_coefs(FIXTURES / fname))
    for z in (0.7, 5.0, -5.0, 30.0, -120.0):
        for eps in (0.0, 3e-12, 8e-11):
            x0 = np.array([[eps * abs(z), 0.0, z]])
            Hk = _as_matrix(o.eval_deriv(x0)[2][0])
            Hf = _fd_hessian(o, np.array([[0.0, 0.0, z]]), 1e-3 * abs(z))
            assert np.max(np.abs(Hk - Hf)) < 1e-7 * np.abs(Hf).max(), (fname, z, eps)
            if fname != "halo":
                assert np.abs(Hf[0, 0] - Hf[1, 1]) > 1e-4 * np.abs(Hf).max()


# ---- extrapolation laws: tables whose continuation beyond the grid is fixed ----
LAW_CASES = ["inner_log", "inner_core", "outer_log", "kepler", "outer_power"]


@pytest.mark.parametrize("name", LAW_CASES)
def test_extrapolation_law_reproduced(name):
    """A monopole table sampled from one of the l=0 extrapolation laws is
    continued by exactly that law beyond the grid (closed forms in
    tests/exact_fields.py): value, force and the radial Hessian
    (D2 - 2D) x x / r^4 + D / r^2, with D, D2 the ln r derivatives.  The
    reference's coefficient (D2 - D) fails the Hessian check; a Lambert-W root
    other than s = v, or a missing s = 2 override, fails the value check."""
    from exact_fields import radial_exact, radial_points, radial_table
    o = OracleMultipole.from_coefs(radial_table(name))
    x = radial_points(name, 300, 1)
    p, F, D = o.eval_deriv(x)
    pe, Fe, De = radial_exact(name, x)
    # relative to |Phi| + |r dPhi/dr| (the log laws pass through Phi = 0)
    scale = np.abs(pe) + np.linalg.norm(Fe, axis=1) * np.linalg.norm(x, axis=1)
    assert np.max(np.abs(p - pe) / scale) < 1e-12
    assert np.max(np.abs(F - Fe)) < 1e-11 * np.abs(Fe).max()
    assert np.max(np.abs(D - De)) < 1e-10 * np.abs(De).max()
    # density: Laplacian / (4 pi G) of the law; 0 beyond r_max (reference :851)
    rho = o.density(x)
    if name.startswith("inner"):
        assert np.max(np.abs(rho + (De[:, 0] + De[:, 1] + De[:, 2]) * INV_4PIG)) \
            < 1e-10 * np.abs(De[:, :3].sum(1) * INV_4PIG).max()
    else:
        assert np.all(rho == 0.0)


def test_law_modes_chosen():
    """The law each exact table selects: log (s = v), the s = 2 core, or the
    Keplerian fallback."""
    from exact_fields import radial_table
    want = {"inner_log": ("inner_mode", "log"), "inner_core": ("inner_mode", "s2"),
            "outer_log": ("outer_mode", "log"), "kepler": ("outer_mode", "kepler"),
            "outer_power": ("outer_mode", "log")}
    for name, (k, v) in want.items():
        assert build_multipole_data(radial_table(name))[k] == v, name


@pytest.mark.parametrize("fname", FILES)
@pytest.mark.parametrize("region", ["inner", "outer"])
def test_calculus_beyond_grid(fname, region):
    """force = -grad Phi and deriv = -grad force by 5-point central
    differences (step 1e-3 r, truncation ~1e-12) below r_min and above r_max,
    where the l=0 laws apply (the interior test above samples r in [0.3, 300]
    only).  Inside r_min the potential is nearly constant (the FIRE table's
    s = 2 core: |Phi| ~ 1e5, |F| ~ 1 at r ~ 1e-4), which bounds the
    differences' round-off; r >= 0.1 r_min keeps it below 1e-7."""
    c = read_coefs(FIXTURES / fname)
    o = OracleMultipole.from_coefs(c)
    rng = np.random.default_rng(8)
    u = unit_vectors(200, 9)
    if region == "inner":
        r = c.R_grid[0] * np.exp(rng.uniform(np.log(0.1), np.log(0.9), 200))
    else:
        r = c.R_grid[-1] * np.exp(rng.uniform(np.log(1.1), np.log(100.0), 200))
    x = u * r[:, None]
    p, F, D = o.eval_deriv(x)
    h = 1e-3 * r[:, None]

    def d5(f, e):
        d = (8 * (f(x + h * e) - f(x - h * e)) - (f(x + 2 * h * e) - f(x - 2 * h * e)))
        return d / (12 * (h[:, 0] if d.ndim == 1 else h))

    g = np.stack([d5(o.potential, e) for e in np.eye(3)], axis=1)
    assert np.max(np.abs(g + F) / np.abs(F).max(1, keepdims=True)) < 1e-7
    cols = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    dF = [d5(lambda y: o.pot_force(y)[1], e) for e in np.eye(3)]
    for slot, (i, j) in enumerate(cols):
        assert np.max(np.abs(dF[j][:, i] - D[:, slot]) / np.abs(D).max(1)) < 1e-7, slot
    # radial: the force is parallel to x
    cross = np.cross(F, x)
    assert np.max(np.linalg.norm(cross, axis=1) / (np.linalg.norm(F, axis=1) * r)) < 1e-12
