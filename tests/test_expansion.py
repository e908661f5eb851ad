"""Pins the coefficient builder (agama_gpu_amd/expansion.py) against what the
mathematics fixes, without Agama (the reference builds these tables with
Agama's C++, gpu_potential.py:688-727, 1935-2052):

* Hernquist, Plummer and spherical Dehnen densities give their closed-form
  Phi_00 and dPhi_00/dr at the nodes, and the error falls at the Simpson
  rule's 4th order when the radial quadrature is refined;
* a manufactured density -- Hernquist plus Gaussian-damped solid harmonics
  g_l(r) B_lm with g_l = (r/s)^l exp(-r^2/s^2), whose potential is exactly
  Phi_H + sum c g_l B_lm -- gives every coefficient column, l up to 8 and
  m of both signs (a wrong (2l+1), normalisation of B_lm or a lost part of
  either Green's integral fails it);
* the King ODE reproduces King's (1966) concentrations c = log10(r_t/r_0),
  and the King table obeys Poisson's equation with the King density;
* the Disk composite (DiskAnsatz + Multipole of the residual density) equals
  the potential of a double-exponential disk from the Hankel integral of
  Kuijken & Gilmore (1989), which a brute-force ring quadrature confirms.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from agama_gpu_amd import expansion as E
from agama_gpu_amd.multipole_prep import G_AGAMA as G
from oracle.oracle import INV_4PIG, OracleMultipole


def _hernquist_rho(M, a):
    def rho(x, y, z):
        r = np.sqrt(x * x + y * y + z * z)
        r = np.where(r < 1e-300, 1e-300, r)
        return M * a / (2.0 * math.pi * r * (r + a) ** 3)
    return rho


def test_hernquist_closed_form_and_order():
    M, a = 1e12, 20.0
    errs = []
    for ff in (8, 16, 40):
        c = E.multipole_from_density(_hernquist_rho(M, a), 0, 0.1, 500.0, 50, fine_factor=ff,
                                     symmetry="spherical")
        r = c.R_grid
        ep = np.max(np.abs(c.phi[:, 0] / (-G * M / (r + a)) - 1))
        ed = np.max(np.abs(c.dphi_dr[:, 0] / (G * M / (r + a) ** 2) - 1))
        errs.append((ep, ed))
    assert errs[2][0] < 1e-11 and errs[2][1] < 1e-10
    # halving the fine step cuts the error by ~2^4
    assert 12 < errs[0][0] / errs[1][0] < 20 and 12 < errs[0][1] / errs[1][1] < 20


@pytest.mark.parametrize("kw,closed", [
    (dict(alpha=2.0, beta=5.0, gamma=0.0, scaleRadius=0.5, mass=1e10),
     lambda r: -G * 1e10 / np.sqrt(r * r + 0.25)),
    (dict(alpha=1.0, beta=4.0, gamma=0.5, scaleRadius=3.0, mass=1e11),
     lambda r: -(G * 1e11 / 3.0) * (1 - (r / (r + 3.0)) ** 1.5) / 1.5),
    (dict(alpha=1.0, beta=4.0, gamma=1.0, scaleRadius=20.0, mass=1e12),
     lambda r: -G * 1e12 / (r + 20.0)),
], ids=["plummer", "dehnen_gamma0.5", "hernquist"])
def test_spheroid_special_cases(kw, closed):
    """Agama's Spheroid reduces to Plummer (2, 5, 0), Dehnen (1, 4, gamma) and
    Hernquist (1, 4, 1); the mass normalisation is the Beta-function integral."""
    c = E.spheroid_coefs(**kw)
    assert c.phi.shape[1] == 1
    r = c.R_grid
    assert np.max(np.abs(c.phi[:, 0] / closed(r) - 1)) < 1e-9


def test_spheroid_mass_with_cutoff():
    """With an exponential cutoff the normalisation integral is numerical: the
    total mass from the table's Keplerian tail equals the requested mass."""
    c = E.spheroid_coefs(mass=5e11, scaleRadius=10.0, alpha=1.0, beta=3.0, gamma=1.0,
                         outerCutoffRadius=100.0, cutoffStrength=2.0, rmax=5e3)
    r = c.R_grid[-1]
    assert -c.phi[-1, 0] * r / G == pytest.approx(5e11, rel=1e-8)


def test_manufactured_solution():
    M, a, s = 1e12, 20.0, 15.0
    terms = [(2, 0, -3e3), (1, 1, 2e3), (2, -2, 1.5e3), (3, 0, 8e2), (4, 4, -5e2), (8, -3, 3e2)]
    labels = E.lm_labels(8)
    rho_h = _hernquist_rho(M, a)

    def g(l, r):
        return (r / s) ** l * np.exp(-(r / s) ** 2)

    def lap_g(l, r):         # g'' + 2 g'/r - l(l+1) g / r^2
        return g(l, r) * (4 * r * r / s ** 4 - (4 * l + 6) / s ** 2)

    cache = {}

    def rho(x, y, z):
        # the solver samples (radii x directions) blocks sharing one set of
        # directions: the angular basis is formed once from the first row
        r = np.sqrt(x * x + y * y + z * z)
        r = np.where(r < 1e-300, 1e-300, r)
        if x.shape[-1] not in cache:
            cache[x.shape[-1]] = E.real_basis(8, z[0] / r[0], np.arctan2(y[0], x[0]))
        B = cache[x.shape[-1]]
        out = rho_h(x, y, z)
        for l, m, cc in terms:
            out = out + cc * lap_g(l, r) * B[labels.index((l, m))][None, :] * INV_4PIG
        return out
    c = E.multipole_from_density(rho, 8, 0.1, 500.0, 50, n_theta=24, n_phi=32, fine_factor=40)
    r = c.R_grid
    ex = np.zeros_like(c.phi)
    ex[:, 0] = -G * M / (r + a)
    for l, m, cc in terms:
        ex[:, labels.index((l, m))] = cc * g(l, r)
    assert np.max(np.abs(c.phi - ex)) < 1e-10 * np.abs(ex).max()
    for l, m, cc in terms:
        j = labels.index((l, m))
        assert np.max(np.abs(c.phi[:, j] - ex[:, j])) < 1e-7 * abs(cc), (l, m)
    dex = np.zeros_like(c.phi)
    dex[:, 0] = G * M / (r + a) ** 2
    for l, m, cc in terms:
        dex[:, labels.index((l, m))] = cc * g(l, r) * (l / r - 2 * r / s ** 2)
    assert np.max(np.abs(c.dphi_dr - dex)) < 1e-9 * np.abs(dex).max()


# King (1966) concentrations log10(r_t / r_0) (tabulated to two decimals)
KING_C = [(3.0, 0.67), (6.0, 1.25), (7.0, 1.53), (9.0, 2.12)]


@pytest.mark.parametrize("W0,c_lit", KING_C)
def test_king_concentration(W0, c_lit):
    _, xi_t = E.king_profile(W0)
    assert math.log10(xi_t) == pytest.approx(c_lit, abs=0.015)


def test_king_table_obeys_poisson():
    """The King table's Laplacian / (4 pi G) (the oracle's density of the
    spline table) equals rho_0 rho(W(r)) / rho(W_0), rho_0 = 9 sigma^2 /
    (4 pi G r_0^2); Keplerian beyond the tidal radius."""
    M, r0, W0 = 1e5, 0.01, 6.0
    c = E.king_coefs(mass=M, scaleRadius=r0, W0=W0, gridSizeR=120)
    sol, xi_t = E.king_profile(W0)
    dWt = sol.sol(xi_t)[1]
    sig2 = G * M / (r0 * xi_t ** 2 * abs(dWt))
    rho0 = 9.0 * sig2 / (4.0 * math.pi * G * r0 * r0)
    o = OracleMultipole.from_coefs(c)
    xi = np.exp(np.linspace(math.log(0.05), math.log(0.8 * xi_t), 40))
    x = np.stack([xi * r0, 0 * xi, 0 * xi], axis=1)
    W = sol.sol(xi)[0]
    rho_ex = rho0 * np.exp(W) * special.gammainc(2.5, W) / (np.exp(W0) * special.gammainc(2.5, W0))
    assert np.max(np.abs(o.density(x) - rho_ex)) < 2e-4 * rho_ex.max()
    rr = np.array([[3 * xi_t * r0, 0, 0]])
    assert o.potential(rr)[0] == pytest.approx(-G * M / (3 * xi_t * r0), rel=1e-12)


def _kuijken_gilmore(R, z, S, Rd, h):
    """Potential of rho = S exp(-R/Rd) exp(-|z|/h) / (2h): the Hankel integral
    -2 pi G S Rd^2 int J0(kR) (1 + k^2 Rd^2)^-3/2 Z(k, z) dk with
    Z = (e^{-k|z|} - k h e^{-|z|/h}) / (1 - k^2 h^2)."""
    def f(k):
        kh = k * h
        if abs(1 - kh) > 1e-6:
            Z = (math.exp(-k * abs(z)) - kh * math.exp(-abs(z) / h)) / (1 - kh * kh)
        else:
            Z = (1 + abs(z) / h) * math.exp(-abs(z) / h) / 2
        return special.j0(k * R) * (1 + k * k * Rd * Rd) ** -1.5 * Z
    zs = [0.0] + list(special.jn_zeros(0, 200) / R)
    v = sum(integrate.quad(f, lo, hi, epsabs=0, epsrel=1e-12)[0] for lo, hi in zip(zs[:-1], zs[1:]))
    v += integrate.quad(f, zs[-1], np.inf, limit=500)[0]
    return -2 * math.pi * G * S * Rd * Rd * v


def _ring_potential(R0, z0, S, Rd, h):
    """The same potential by direct quadrature over rings (complete elliptic K)."""
    def inner(Rp):
        def f(zp):
            d2 = (R0 + Rp) ** 2 + (z0 - zp) ** 2
            return math.exp(-abs(zp) / h) / (2 * h) * 4 * Rp * special.ellipk(4 * R0 * Rp / d2) \
                / math.sqrt(d2)
        edges = sorted({-60 * h, 0.0, z0, 60 * h})
        return S * math.exp(-Rp / Rd) * sum(
            integrate.quad(f, lo, hi, limit=200, epsabs=0, epsrel=1e-10)[0]
            for lo, hi in zip(edges[:-1], edges[1:]) if hi > lo)
    tot = sum(integrate.quad(inner, lo, hi, limit=400, epsabs=0, epsrel=1e-9)[0]
              for lo, hi in ((0, R0), (R0, 3 * R0), (3 * R0, 60 * Rd)))
    return -G * tot


def test_disk_composite_matches_exact_potential():
    from oracle import analytic_oracle as A
    S, Rd, h = 8e8, 3.0, 0.3
    ansatz, coefs = E.disk_potential_parts(surfaceDensity=S, scaleRadius=Rd, scaleHeight=h)
    assert coefs.lmax == 32 and coefs.symmetry == "Axisymmetric"
    o = OracleMultipole.from_coefs(coefs)
    pts = np.array([[1.0, 0, 0.0], [3.0, 0, 0.0], [8.0, 0, 0.0], [3.0, 0, 0.3], [3.0, 0, 1.0],
                    [15.0, 0, 0.0], [20.0, 0, 5.0], [0.5, 0, 0.1]])
    P = A.disk_ansatz(pts, S, Rd, h, 0.0)[0] + o.potential(pts)
    ref = np.array([_kuijken_gilmore(R, z, S, Rd, h) for R, _, z in pts])
    assert np.max(np.abs(P - ref) / np.abs(ref)) < 2e-5
    assert _ring_potential(3.0, 0.0, S, Rd, h) == pytest.approx(ref[1], rel=1e-9)


def test_disk_mass_normalisation():
    """``mass`` in place of ``surfaceDensity``: M = 2 pi Sigma Rd^2 without an
    inner cutoff; the residual's Keplerian tail carries the disk's whole mass
    (the ansatz carries none)."""
    an, c = E.disk_potential_parts(mass=6e10, scaleRadius=3.0, scaleHeight=-0.3, lmax=16)
    assert an["surfaceDensity"] == pytest.approx(6e10 / (2 * math.pi * 9.0), rel=1e-14)
    assert -c.phi[-1, 0] * c.R_grid[-1] / G == pytest.approx(6e10, rel=1e-6)
    an, _ = E.disk_potential_parts(mass=6e10, scaleRadius=3.0, scaleHeight=0.3,
                                   innerCutoffRadius=1.0, lmax=8)
    val = integrate.quad(lambda R: R * math.exp(-R / 3.0 - 1.0 / R), 0, np.inf)[0]
    assert an["surfaceDensity"] == pytest.approx(6e10 / (2 * math.pi * val), rel=1e-10)


def test_factory_builds_agama_routed_types():
    """PotentialGPU(type=...) accepts Spheroid, King and Disk (built here, not
    by Agama): the construction of their tables runs on the CPU; evaluation
    is covered by the GPU tests."""
    from agama_gpu_amd import factory
    tm = factory.type_map()
    for t in ("spheroid", "king", "disk", "dehnen"):
        assert callable(tm[t])
    with pytest.raises(ValueError, match="W0"):
        factory._king(mass=1.0)
    with pytest.raises(NotImplementedError, match="sersicIndex"):
        E.disk_potential_parts(surfaceDensity=1.0, sersicIndex=2.0)
