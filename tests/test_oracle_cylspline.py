"""Pins the CylSpline oracle (oracle/cylspline_oracle.py) without Agama:
its spline solver against scipy, the product's table builder against it,
node exactness on the fixture grid, calculus identities inside the grid, and
harmonicity of the outer PowerLaw expansion (zero Laplacian)."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from conftest import FIXTURES
from oracle.cylspline_oracle import OracleCylSpline, spline_slopes
from agama_gpu_amd.coefs import read_coefs
from agama_gpu_amd.cylspline_prep import build_cylspline_data

FILE = FIXTURES / "600.bar.none_8.coef_cylsp_DR"


@pytest.fixture(scope="module")
def coefs():
    return read_coefs(FILE)


@pytest.fixture(scope="module")
def orc(coefs):
    return OracleCylSpline(coefs)


def test_spline_slopes_match_scipy():
    rng = np.random.default_rng(0)
    x = np.cumsum(rng.uniform(0.1, 1.0, 30))
    f = np.sin(x) + 0.1 * x ** 2
    assert np.allclose(spline_slopes(x, f), CubicSpline(x, f, bc_type="natural")(x, 1), rtol=1e-12, atol=1e-12)
    assert np.allclose(spline_slopes(x, f, 0.0),
                       CubicSpline(x, f, bc_type=((1, 0.0), "natural"))(x, 1), rtol=1e-12, atol=1e-12)


def test_product_table_builder_matches(coefs, orc):
    d = build_cylspline_data(coefs)
    assert d["Rscale"] == pytest.approx(orc.Rscale, rel=1e-12)
    assert np.allclose(d["W_outer"], orc.W, rtol=1e-9, atol=1e-9 * np.abs(orc.W).max())
    for i, m in enumerate(d["m_arr"]):
        f, fx, fy, fxy = orc.harm[int(m)]
        got = d["node_arr"][i]
        for k, ref in enumerate((f, fx, fy, fxy)):
            assert np.max(np.abs(got[..., k] - ref)) <= 1e-10 * max(np.abs(ref).max(), 1e-300), (m, k)


def test_node_exactness(coefs, orc):
    R, z = coefs.R_grid, coefs.z_grid
    rng = np.random.default_rng(1)
    iR = rng.integers(1, len(R) - 1, 40)
    iz = rng.integers(1, len(z) - 1, 40)
    ph = rng.uniform(0, 2 * np.pi, 40)
    x = np.stack([R[iR] * np.cos(ph), R[iR] * np.sin(ph), z[iz]], axis=1)
    ex = np.zeros(40)
    for m, t in coefs.phi.items():
        T = np.cos(m * ph) if m >= 0 else np.sin(-m * ph)
        ex += t[iR, iz] * T
    p = orc.eval(x)[0]
    assert np.max(np.abs(p - ex) / np.abs(ex)) < 1e-12


def _grid_points(orc, n, seed, inside=True):
    rng = np.random.default_rng(seed)
    if inside:
        R = orc.Rscale * np.sinh(rng.uniform(orc.lR[0], orc.lR[-1], n))
        z = orc.Rscale * np.sinh(rng.uniform(orc.lz[1], orc.lz[-2], n))
    else:
        r = np.exp(rng.uniform(np.log(720.0), np.log(3000.0), n))
        ct = rng.uniform(-1, 1, n)
        R, z = r * np.sqrt(1 - ct * ct), r * ct
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.stack([R * np.cos(ph), R * np.sin(ph), z], axis=1)


def test_calculus_inside(orc):
    x = _grid_points(orc, 300, 2)
    P, g, H, outer = orc.eval(x)
    assert not outer.any()
    step = 1e-6 * np.linalg.norm(x, axis=1)[:, None]
    e = np.eye(3)
    fd = np.stack([(orc.eval(x + step * e[k])[0] - orc.eval(x - step * e[k])[0]) / (2 * step[:, 0])
                   for k in range(3)], axis=1)
    assert np.max(np.abs(fd - g)) <= 1e-6 * np.abs(g).max()
    gp = [orc.eval(x + step * e[k])[1] for k in range(3)]
    gm = [orc.eval(x - step * e[k])[1] for k in range(3)]
    err = []
    for s, (i, j) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]):
        fdh = (gp[j][:, i] - gm[j][:, i]) / (2 * step[:, 0])
        err.append(np.abs(fdh - H[:, s]))
    err = np.max(np.stack(err, axis=1), axis=1)
    # a step may cross a spline cell boundary where second derivatives jump
    assert np.mean(err <= 1e-4 * np.abs(H).max()) > 0.97


def test_outer_expansion_is_harmonic(orc):
    x = _grid_points(orc, 200, 3, inside=False)
    P, g, H, outer = orc.eval(x)
    assert outer.all()
    lap = H[:, 0] + H[:, 1] + H[:, 2]
    assert np.max(np.abs(lap)) <= 1e-6 * np.abs(H).max()
    step = 1e-6 * np.linalg.norm(x, axis=1)[:, None]
    e = np.eye(3)
    fd = np.stack([(orc.eval(x + step * e[k])[0] - orc.eval(x - step * e[k])[0]) / (2 * step[:, 0])
                   for k in range(3)], axis=1)
    assert np.max(np.abs(fd - g)) <= 1e-6 * np.abs(g).max()


# ---- the outer PowerLaw fit against a field the mathematics fixes -------------
def test_outer_fit_recovers_exterior_field():
    """Tables sampled from an exterior field of solid harmonics (monopole,
    dipole, quadrupole m = 0, +-2, octupole; tests/exact_fields.py, which
    derives each W_lm from Agama's normalisation by hand): the boundary fit of
    the product (its own Legendre recurrence) and of the oracle (scipy
    sph_harm_y) both recover every W_lm to lstsq precision and give zero for
    the absent terms.  The reference's recurrence (0 / P_(l-3) in place of
    P_(l-2), _cylspl_potential_kernel.cu:205-223) fails this; so would a wrong
    normalisation, a dropped Condon-Shortley sign or a transposed (l, m)."""
    from exact_fields import exterior_W, exterior_coefs
    from agama_gpu_amd.cylspline_prep import outer_asymptote
    c = exterior_coefs()
    W, r0 = outer_asymptote(c.R_grid, c.z_grid, c.phi, c.mmax)
    assert r0 == 60.0
    We = exterior_W(r0)
    assert np.max(np.abs(W - We)) < 1e-12 * np.abs(We).max()
    Wo, r0o = OracleCylSpline._fit(c.R_grid, c.z_grid, c.phi, c.mmax)
    assert r0o == r0 and np.max(np.abs(Wo - We)) < 1e-12 * np.abs(We).max()


def test_rscale_and_outer_values_follow():
    """With the fitted W: Rscale = -Mtot / Phi0 with Mtot = -W_00 r0 2 sqrt(pi)
    and Phi0 the m = 0 node at R = 0, iz = (nz+1)//2 (gpu_potential.py:
    1188-1207); outside the grid box the oracle equals the generating field,
    and across the box's boundary nodes the outer expansion continues the
    inner spline (both equal the node value there)."""
    from exact_fields import exterior_W, exterior_coefs, exterior_phi
    c = exterior_coefs()
    d = build_cylspline_data(c)
    o = OracleCylSpline(c)
    We = exterior_W(60.0)
    Phi0 = c.phi[0][0, (len(c.z_grid) + 1) // 2]
    want = -(-We[0] * 60.0 * 2 * np.sqrt(np.pi)) / Phi0
    assert d["Rscale"] == pytest.approx(want, rel=1e-12) and o.Rscale == pytest.approx(want, rel=1e-12)
    rng = np.random.default_rng(4)
    n = 400
    r = np.exp(rng.uniform(np.log(110.0), np.log(2000.0), n))
    ct = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    x = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)
    P, g, H, outer = o.eval(x)
    assert outer.all()
    ex = exterior_phi(x[:, 0], x[:, 1], x[:, 2])
    assert np.max(np.abs(P - ex) / np.abs(ex)) < 1e-12
    # gradient of the closed form by complex-step differentiation
    ge = np.stack([np.imag(exterior_phi(*(x + 1e-30j * e).T)) / 1e-30 for e in np.eye(3)], axis=1)
    assert np.max(np.abs(g - ge)) < 1e-11 * np.abs(ge).max()
    # boundary nodes (top edge and outer R edge), just inside vs just outside
    R, z = c.R_grid, c.z_grid
    pts = [(R[i], z[-1]) for i in range(1, len(R) - 1)] + [(R[-1], z[j]) for j in range(len(z))]
    pts = np.array(pts)
    ph = rng.uniform(0, 2 * np.pi, len(pts))
    base = np.stack([pts[:, 0] * np.cos(ph), pts[:, 0] * np.sin(ph), pts[:, 1]], axis=1)
    for eps in (1e-9,):
        pin, _, _, oin = o.eval(base * (1 - eps))
        pout, _, _, oout = o.eval(base * (1 + eps))
        assert not oin.any() and oout.all()
        assert np.max(np.abs(pin - pout) / np.abs(pout)) < 1e-8


# ---- exactly on the z axis ---------------------------------------------------
@pytest.mark.parametrize("kind", ["log", "linear"])
def test_axis_hessian_is_the_limit(kind):
    """On the axis the Hessian equals central differences of the off-axis
    gradient across it, Richardson-extrapolated from steps of 1e-5 and
    5e-6 Rscale along x, y, z (the spline's |R|^3 terms make one difference
    first-order accurate): H_xx and H_yy differ by the m = 2 terms and H_xy
    is the m = -2 term, which the reference's R = 0 branch (H_xx = H_yy =
    d2Phi/dR2 at phi = 0, H_xy = 0; _cylspl_potential_kernel.cu:855-863)
    misses."""
    from exact_fields import axis_tables
    o = OracleCylSpline(axis_tables()[kind])
    assert o.log_scaling == (kind == "log")
    z = np.array([-60.0, -7.5, -0.3, 0.3, 1.0, 4.0, 25.0, 70.0])
    on = np.stack([0 * z, 0 * z, z], axis=1)
    P, g, H, outer = o.eval(on)
    assert not outer.any()

    def diffs(d):
        gp = [o.eval(on + d * e)[1] for e in np.eye(3)]
        gm = [o.eval(on - d * e)[1] for e in np.eye(3)]
        # slots xx, yy, zz, xy, yz, xz as (component, direction)
        return np.stack([(gp[j][:, i] - gm[j][:, i]) / (2 * d)
                         for i, j in [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]], axis=1)

    d = 1e-5 * o.Rscale
    fd = 2 * diffs(d / 2) - diffs(d)
    scale = np.abs(H).max()
    assert np.max(np.abs(fd - H)) <= 1e-8 * scale
    # the m = 2 terms are not negligible here: H_xx != H_yy, H_xy != 0
    assert np.abs(H[:, 0] - H[:, 1]).max() > 1e-3 * scale
    assert np.abs(H[:, 3]).max() > 1e-3 * scale


def test_axis_hessian_is_the_direction_average(orc):
    """For a table with every harmonic (the bar fixture: m = -8..8, zero on the
    axis for m != 0), the Hessian of the m = +-1 and |m| >= 3 terms has no
    limit at the axis -- it depends on the direction of approach -- and the
    axis value is its average over directions: the mean of the off-axis
    Hessian over 64 equally spaced azimuths, Richardson-extrapolated to R = 0
    from R = 1e-6 and 5e-7 Rscale (the mean is R-linear at first order)."""
    z = np.array([-40.0, -3.0, 0.2, 2.0, 9.0, 150.0])
    on = np.stack([0 * z, 0 * z, z], axis=1)
    P, g, H, outer = orc.eval(on)
    assert not outer.any()
    ph = 2 * np.pi * np.arange(64) / 64

    def ring(R):
        x = np.stack([R * np.tile(np.cos(ph), len(z)), R * np.tile(np.sin(ph), len(z)),
                      np.repeat(z, 64)], axis=1)
        return orc.eval(x)[2].reshape(len(z), 64, 6)

    R = 1e-6 * orc.Rscale
    Hr = ring(R)
    scale = np.abs(H).max()
    assert np.max(np.abs(2 * ring(R / 2).mean(1) - Hr.mean(1) - H)) <= 1e-9 * scale
    # the directions do differ (the average is not a single direction's value)
    assert np.abs(Hr[:, :, 0] - Hr[:, :1, 0]).max() > 1e-4 * scale
    assert np.abs(H[:, 0] - H[:, 1]).max() > 1e-4 * scale
