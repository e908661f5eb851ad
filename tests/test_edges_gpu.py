"""GPU edge cases: the grid's own radii, the extrapolation laws against closed
forms, the z-axis band, bad positions, and the time-dependent wrapper modes.

* Node radii, r_min, r_max and r_max (1 +- 1e-15) on the reference's two
  Multipole fixtures, every kernel path (reference tests/test_gpu_audit.py:
  124-169, tests/audit_gpu_vs_cpu.py:102-160; k clamping at
  _multipole_potential_kernel.cu:195-197, 216): at interior nodes the result
  equals the file's Phi_lm(r_i) summed over scipy's spherical harmonics; at
  the grid edges, where the branch (spline or l=0 law) turns on the last bit
  of ln r, it equals the oracle's value of one of the two branches.
* Monopole tables sampled from the extrapolation laws (tests/exact_fields.py)
  reproduce them beyond the grid on the GPU: value, force, Hessian.
* The CylSpline exterior-field table: outside the grid box the GPU equals the
  generating solid-harmonic field (value 1e-12, force 1e-11, Hessian against
  differences of its complex-step gradient 1e-7).
* The approach to the z axis at both poles: phi and force continuous with
  the on-axis values.
* NaN positions give NaN for every type; |x| = 1e200 (r^2 overflows) gives
  the Multipole outer law's limit 0, as the reference does; NFW at the exact
  origin gives Phi = -GM/r_s and zero force.
* Shifted with a (T,7) Hermite centre trajectory and Scaled with (T,2) and
  (T,3) schedules (gpu_potential.py:1693-1736, 1767-1876), inside the time
  range and extrapolated.
"""
import math

import numpy as np
import pytest
import torch

from conftest import FIXTURES
from oracle import analytic_oracle as A

pytestmark = pytest.mark.gpu
DEV = "cuda"
FILES = ["600.dark.none_8.coef_mul_DR", "100.LMC.none_8.coef_mult"]


def unit_vectors(n, seed):
    u = np.random.default_rng(seed).normal(size=(n, 3))
    return u / np.linalg.norm(u, axis=1)[:, None]


def basis(labels, u):
    from scipy.special import sph_harm_y
    th = np.arccos(u[:, 2])
    ph = np.arctan2(u[:, 1], u[:, 0])
    cols = []
    for l, m in labels:
        am = abs(m)
        Pt = np.real(sph_harm_y(l, am, th, ph) * np.exp(-1j * am * ph))
        mul = 2 * np.sqrt(np.pi) if m == 0 else 2 * np.sqrt(2 * np.pi)
        cols.append(mul * Pt * (np.cos(am * ph) if m >= 0 else np.sin(am * ph)))
    return np.asarray(cols).T


@pytest.fixture(params=[1, 2, 3, 4], ids=["lane", "binned", "lds", "sorted"])
def mp_path(request):
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_MP_PATH, request.param)
    yield request.param
    _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)


@pytest.mark.parametrize("fname", FILES)
def test_node_radii(fname, mp_path):
    from agama_gpu_amd import MultipolePotentialGPU, read_coefs
    from oracle.oracle import OracleMultipole
    print(f"node radii {fname} path={mp_path}", flush=True)
    c = read_coefs(FIXTURES / fname)
    gpu = MultipolePotentialGPU(c)
    orc = OracleMultipole.from_coefs(c)
    u = unit_vectors(64, 5)
    B = basis(c.lm_labels, u)
    nR = len(c.R_grid)
    # interior nodes: exactness (the files are log-uniform to ~6e-13, which
    # bounds it), and the oracle
    x = np.concatenate([u * c.R_grid[i] for i in range(1, nR - 1)])
    # 4096+ points so that the LDS / sorted paths take them
    reps = 4096 // len(x) + 1
    xt = torch.as_tensor(np.tile(x, (reps, 1)), device=DEV)
    p, F = [t.cpu().numpy()[:len(x)] for t in gpu.eval(xt, pot=True, acc=True)]
    po, Fo = orc.pot_force(x)
    assert np.max(np.abs(p - po) / np.abs(po)) < 1e-12
    assert np.max(np.abs(F - Fo)) < 1e-11 * np.abs(Fo).max()
    for i in range(1, nR - 1):
        sl = slice((i - 1) * len(u), i * len(u))
        ex = B @ c.phi[i]
        assert np.max(np.abs(p[sl] - ex)) < 1e-10 * np.abs(ex).max(), i
        dr = -(F[sl] * u).sum(1)
        exd = B @ c.dphi_dr[i]
        assert np.max(np.abs(dr - exd)) < 1e-8 * np.abs(exd).max(), i
    # grid edges: one of the two branches' values, nothing else (no
    # out-of-range slab read at k = n_int)
    r0, rN = c.R_grid[0], c.R_grid[-1]
    for r, f_in, f_out in ((r0, 1 + 1e-12, 1 - 1e-12), (rN, 1 - 1e-12, 1 + 1e-12),
                           (rN * (1 + 1e-15), 1 - 1e-12, 1 + 1e-12),
                           (rN * (1 - 1e-15), 1 - 1e-12, 1 + 1e-12)):
        xe = u * r
        xt = torch.as_tensor(np.tile(xe, (4096 // len(xe) + 1, 1)), device=DEV)
        pg, Fg = [t.cpu().numpy()[:len(xe)] for t in gpu.eval(xt, pot=True, acc=True)]
        pa, Fa = orc.pot_force(xe * f_in)      # the spline branch's limit
        pb, Fb = orc.pot_force(xe * f_out)     # the l=0 law's limit
        ea = np.abs(pg - pa) / np.abs(pa)
        eb = np.abs(pg - pb) / np.abs(pb)
        assert np.all(np.minimum(ea, eb) < 1e-10), r
        fa = np.abs(Fg - Fa).max(1) / np.abs(Fa).max()
        fb = np.abs(Fg - Fb).max(1) / np.abs(Fb).max()
        assert np.all(np.minimum(fa, fb) < 1e-8), r


LAW_CASES = ["inner_log", "inner_core", "outer_log", "kepler", "outer_power"]


@pytest.mark.parametrize("name", LAW_CASES)
def test_extrapolation_laws_closed_form(name, mp_path):
    from exact_fields import radial_exact, radial_points, radial_table
    from agama_gpu_amd import MultipolePotentialGPU
    print(f"law {name} path={mp_path}", flush=True)
    gpu = MultipolePotentialGPU(radial_table(name))
    x = radial_points(name, 5000, 2)
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(torch.as_tensor(x, device=DEV))]
    pe, Fe, De = radial_exact(name, x)
    scale = np.abs(pe) + np.linalg.norm(Fe, axis=1) * np.linalg.norm(x, axis=1)
    assert np.max(np.abs(p - pe) / scale) < 1e-12
    assert np.max(np.abs(F - Fe)) < 1e-11 * np.abs(Fe).max()
    assert np.max(np.abs(D - De)) < 1e-10 * np.abs(De).max()


def test_cylspline_exterior_field():
    from exact_fields import exterior_coefs, exterior_phi
    from agama_gpu_amd import CylSplinePotentialGPU
    from oracle.cylspline_oracle import OracleCylSpline
    print("cylspline exterior field", flush=True)
    c = exterior_coefs()
    gpu = CylSplinePotentialGPU(c)
    rng = np.random.default_rng(6)
    n = 20000
    r = np.exp(rng.uniform(np.log(110.0), np.log(3000.0), n))
    ct = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    x = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(torch.as_tensor(x, device=DEV))]
    ex = exterior_phi(*x.T)
    assert np.max(np.abs(p - ex) / np.abs(ex)) < 1e-12

    def grad(y):   # complex-step gradient of the closed form
        return np.stack([np.imag(exterior_phi(*(y + 1e-30j * e).T)) / 1e-30 for e in np.eye(3)],
                        axis=1)
    ge = grad(x)
    assert np.max(np.abs(F + ge)) < 1e-11 * np.abs(ge).max()
    h = 1e-5 * r[:, None]
    slots = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    gp = [grad(x + h * e) for e in np.eye(3)]
    gm = [grad(x - h * e) for e in np.eye(3)]
    He = np.stack([(gp[j][:, i] - gm[j][:, i]) / (2 * h[:, 0]) for i, j in slots], axis=1)
    assert np.max(np.abs(D + He)) < 1e-7 * np.abs(He).max()
    # inside the grid: the oracle on the same table
    orc = OracleCylSpline(c)
    R = orc.Rscale * np.sinh(rng.uniform(orc.lR[0], orc.lR[-1], n))
    z = orc.Rscale * np.sinh(rng.uniform(orc.lz[0], orc.lz[-1], n))
    xi = np.stack([R * np.cos(ph), R * np.sin(ph), z], axis=1)
    p, F = [t.cpu().numpy() for t in gpu.eval(torch.as_tensor(xi, device=DEV), pot=True, acc=True)]
    Po, go, _, outer = orc.eval(xi)
    assert not outer.any()
    assert np.max(np.abs(p - Po) / np.abs(Po)) < 1e-12
    assert np.max(np.abs(F + go)) < 1e-11 * np.abs(go).max()


@pytest.mark.parametrize("fname", FILES)
def test_near_axis_force(fname, mp_path):
    """Points at R = eps r, eps in [1e-12, 1e-6], approaching the z axis at
    both poles and several radii: phi and force are continuous with the
    on-axis values (the m = 1 terms' Y / sin(theta) limit is -dY/dtheta at the
    south pole; reference tests/test_gpu_audit.py:176-223), and on the axis
    the GPU equals the oracle."""
    from agama_gpu_amd import MultipolePotentialGPU, read_coefs
    from oracle.oracle import OracleMultipole
    print(f"near axis {fname} path={mp_path}", flush=True)
    c = read_coefs(FIXTURES / fname)
    gpu = MultipolePotentialGPU(c)
    rng = np.random.default_rng(3)
    n = 8192
    r = np.exp(rng.uniform(np.log(0.5), np.log(200.0), n))
    eps = 10 ** rng.uniform(-12, -6, n)
    sgn = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    ph = rng.uniform(0, 2 * np.pi, n)
    x = np.stack([eps * r * np.cos(ph), eps * r * np.sin(ph), sgn * r * np.sqrt(1 - eps ** 2)], axis=1)
    on = np.stack([0 * r, 0 * r, sgn * r], axis=1)
    p, F = [t.cpu().numpy() for t in gpu.eval(torch.as_tensor(x, device=DEV), pot=True, acc=True)]
    po, Fo = [t.cpu().numpy() for t in gpu.eval(torch.as_tensor(on, device=DEV), pot=True, acc=True)]
    assert np.isfinite(p).all() and np.isfinite(F).all()
    # phi varies by O(eps) off the axis (the m = 1 terms)
    assert np.max(np.abs(p - po) / np.abs(po)) < 1e-6
    assert np.max(np.abs(F - Fo).max(1) / np.abs(Fo).max(1)) < 1e-5
    orc = OracleMultipole.from_coefs(c)
    pr, Fr = orc.pot_force(on)
    assert np.max(np.abs(po - pr) / np.abs(pr)) < 1e-12
    assert np.max(np.abs(Fo - Fr)) < 1e-11 * np.abs(Fr).max()
    # Hessian: on the axis, and within sin(theta) < 1e-10 of it, the regular
    # limit (equal to the oracle's closed form, itself pinned by differences of
    # the force); between 1e-10 and 1e-6 the spherical chain rule, whose
    # cancelling 1/sin(theta) terms leave an error of order 1e-16 / sin^2
    # (theta) of max|H|: finite, and that close to the axis value
    D = gpu.evalDeriv(torch.as_tensor(x, device=DEV))[2].cpu().numpy()
    Do = gpu.evalDeriv(torch.as_tensor(on, device=DEV))[2].cpu().numpy()
    Dr = orc.eval_deriv(on)[2]
    assert np.isfinite(D).all()
    assert np.max(np.abs(Do - Dr)) < 1e-10 * np.abs(Dr).max()
    rel = np.abs(D - Do).max(1) / np.abs(Do).max(1)
    pole = eps < 1e-10
    print(f"  pole rows max rel {rel[pole].max():.2e}, band max rel*eps^2 "
          f"{(rel * eps ** 2)[~pole].max():.2e}, band max rel/eps {(rel / eps)[~pole].max():.2e}",
          flush=True)
    assert rel[pole].max() < 1e-8
    assert np.all(rel[~pole] < 1e-4 + 1e-15 / eps[~pole] ** 2)      # measured: 7e-17 / eps^2


def _bad_points():
    x = np.array([[np.nan, 1.0, 2.0], [1.0, np.nan, 2.0], [1.0, 2.0, np.nan],
                  [1e200, 0.0, 0.0], [0.0, -1e200, 3.0], [1.0, 1.0, 1e170]])
    return x, np.array([True, True, True, False, False, False])


def test_bad_positions_multipole(mp_path):
    """NaN in any coordinate gives NaN phi and force (the reference: NaN
    reaches its spline); r^2 = inf gives 0 (the reference's outer law at
    r = inf).  Padded to 8192 points so every path's traversal sees them."""
    from agama_gpu_amd import MultipolePotentialGPU, read_coefs
    from agama_gpu_amd.workloads import hernquist_points
    print(f"bad positions path={mp_path}", flush=True)
    gpu = MultipolePotentialGPU(read_coefs(FIXTURES / FILES[0]))
    bad, isnan = _bad_points()
    good = hernquist_points(8192, 3)
    x = np.concatenate([good, bad, good[:100]])
    xt = torch.as_tensor(x, device=DEV)
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(xt)]
    p1, F1 = [t.cpu().numpy() for t in gpu.eval(xt, pot=True, acc=True)]
    gt = torch.as_tensor(good, device=DEV)
    ref1 = [t.cpu().numpy() for t in gpu.eval(gt, pot=True, acc=True)]
    ref2 = [t.cpu().numpy() for t in gpu.evalDeriv(gt)][:2]
    b = slice(8192, 8192 + len(bad))
    for pp, FF, (rp, rF) in ((p, F, ref2), (p1, F1, ref1)):
        assert np.all(np.isnan(pp[b][isnan])) and np.all(np.isnan(FF[b][isnan]).any(1))
        assert np.all(pp[b][~isnan] == 0.0) and np.all(FF[b][~isnan] == 0.0)
        # the good points around them are unaffected (same kernel, same bits)
        assert np.array_equal(pp[:8192], rp) and np.array_equal(FF[:8192], rF)
    assert np.all(D[b][~isnan] == 0.0) and np.all(np.isnan(D[b][isnan]).any(1))
    rho = gpu.density(xt).cpu().numpy()
    assert np.all(np.isnan(rho[b][isnan])) and np.all(rho[b][~isnan] == 0.0)


@pytest.mark.parametrize("name,kw", [
    ("NFW", dict(mass=1e12, scaleRadius=20.0)),
    ("Plummer", dict(mass=1e10, scaleRadius=0.5)),
    ("Hernquist", dict(mass=5e11, scaleRadius=10.0)),
    ("Dehnen", dict(mass=1e11, scaleRadius=3.0, gamma=0.5)),
    ("Isochrone", dict(mass=1e11, scaleRadius=2.0)),
    ("MiyamotoNagai", dict(mass=6e10, scaleRadius=3.0, scaleHeight=0.3)),
    ("LogHalo", dict(velocity=200.0, coreRadius=1.0, axisRatioY=0.9, axisRatioZ=0.7)),
    ("DiskAnsatz", dict(surfaceDensity=8e8, scaleRadius=3.0, scaleHeight=0.4)),
    ("CylSpline", None), ("Composite", None)])
def test_bad_positions_nan(name, kw):
    """NaN positions give NaN for the analytic types, CylSpline and a fused
    composite (the fmax clamps that used to map them to the centre or the
    axis are gone)."""
    from agama_gpu_amd import CylSplinePotentialGPU, MultipolePotentialGPU, PotentialGPU, read_coefs
    if name == "CylSpline":
        pot = CylSplinePotentialGPU(read_coefs(FIXTURES / "600.bar.none_8.coef_cylsp_DR"))
    elif name == "Composite":
        pot = (MultipolePotentialGPU(read_coefs(FIXTURES / FILES[1])) +
               CylSplinePotentialGPU(read_coefs(FIXTURES / "600.bar.none_8.coef_cylsp_DR")) +
               PotentialGPU(type="NFW", mass=1e12, scaleRadius=20.0))
    else:
        pot = PotentialGPU(type=name, **kw)
    bad, isnan = _bad_points()
    x = np.concatenate([np.array([[3.0, -2.0, 1.0]]), bad[isnan]])
    p, F, D = [t.cpu().numpy() for t in pot.evalDeriv(torch.as_tensor(x, device=DEV))]
    assert np.isfinite(p[0]) and np.isfinite(F[0]).all()
    assert np.all(np.isnan(p[1:])) and np.all(np.isnan(F[1:]).any(1)) and np.all(np.isnan(D[1:]).any(1))


def test_nfw_origin():
    """NFW at r = 0: Phi = -GM/r_s (the limit) and zero force (by symmetry),
    where the closed-form gradient (reference _analytic_potentials.py:
    200-214) gives NaN; near the origin the small-r series and the closed
    form agree with mpmath to 1e-13 (the closed form alone loses digits as
    r_s/r by cancellation)."""
    import mpmath as mp
    from agama_gpu_amd import PotentialGPU
    M, rs = 1e12, 20.0
    GM = A.G * M
    pot = PotentialGPU(type="NFW", mass=M, scaleRadius=rs)
    r = np.array([0.0, 1e-9, 1e-6, 1e-3, 0.1, 1.2, 1.3, 5.0, 50.0])
    x = np.stack([r, 0 * r, 0 * r], axis=1)
    p, F, D = [t.cpu().numpy() for t in pot.evalDeriv(torch.as_tensor(x, device=DEV))]
    assert p[0] == pytest.approx(-GM / rs, rel=1e-15) and np.all(F[0] == 0.0)
    mp.mp.dps = 40

    def phi(rr):
        return -GM * mp.log1p(rr / rs) / rr
    for i in range(1, len(r)):
        rr = mp.mpf(r[i])
        assert abs(p[i] - float(phi(rr))) < 1e-14 * GM / rs
        d1 = float(mp.diff(phi, rr))
        d2 = float(mp.diff(phi, rr, 2))
        assert abs(-F[i, 0] - d1) < 1e-13 * abs(d1), r[i]
        assert abs(-D[i, 0] - d2) < 1e-12 * abs(d2), r[i]
        assert abs(-D[i, 1] - d1 / r[i]) < 1e-13 * abs(d1 / r[i]), r[i]


def test_time_dependent_wrappers():
    """Shifted with a (T,7) centre (Hermite through positions and
    velocities) and Scaled with (T,2) [t, scale] and (T,3) [t, ampl, scale]
    schedules (not-a-knot splines), each linearly extrapolated outside the
    knots, against the oracle at the scipy-evaluated centre / scale / ampl."""
    from scipy.interpolate import CubicHermiteSpline, CubicSpline
    from agama_gpu_amd import MultipolePotentialGPU, PotentialGPU, ScaledPotentialGPU, \
        ShiftedPotentialGPU, read_coefs
    from agama_gpu_amd.workloads import hernquist_points
    from oracle.oracle import OracleMultipole
    x = hernquist_points(20000, 9)
    xt = torch.as_tensor(x, device=DEV)
    nfw = PotentialGPU(type="NFW", mass=1e12, scaleRadius=20.0)
    T = np.linspace(-2.0, 6.0, 9)
    pos = np.stack([30 * np.cos(T / 2), 20 * np.sin(T / 3), 4.0 * T], axis=1)
    vel = np.stack([-15 * np.sin(T / 2), 20 / 3 * np.cos(T / 3), 4.0 + 0 * T], axis=1) * 1.3
    traj = np.concatenate([T[:, None], pos, vel], axis=1)
    sh = ShiftedPotentialGPU(nfw, center=traj[::-1])          # unsorted input is sorted
    spl = CubicHermiteSpline(T, pos, vel)
    for t in (0.37, 5.99, -3.5, 8.25):
        if t < T[0]:
            c = spl(T[0]) + spl(T[0], 1) * (t - T[0])
        elif t > T[-1]:
            c = spl(T[-1]) + spl(T[-1], 1) * (t - T[-1])
        else:
            c = spl(t)
        ref = A.transformed(lambda y: A.nfw(y, 1e12, 20.0), x, tau=c)
        p, F, D = [a.cpu().numpy() for a in sh.evalDeriv(xt, t=t)]
        assert np.max(np.abs(p - ref[0]) / np.abs(ref[0])) < 1e-11, t
        assert np.max(np.abs(F + ref[1])) < 1e-11 * np.abs(ref[1]).max(), t
        assert np.max(np.abs(D + ref[2])) < 1e-11 * np.abs(ref[2]).max(), t

    coefs = read_coefs(FIXTURES / FILES[1])
    mp = MultipolePotentialGPU(coefs)
    orc = OracleMultipole.from_coefs(coefs)

    def mp_ref(y):
        p, F, D = orc.eval_deriv(y)
        return p, -F, -D, orc.density(y)
    Ts = np.array([0.0, 1.0, 2.5, 4.0, 5.0])
    scales = np.array([1.0, 1.3, 1.1, 0.8, 0.9])
    ampls = np.array([1.0, 0.7, 0.9, 1.2, 1.5])
    for cols in (2, 3):
        sched = np.stack([Ts, ampls, scales], axis=1) if cols == 3 else np.stack([Ts, scales], axis=1)
        sc = ScaledPotentialGPU(mp, scale=sched, ampl=0.6)
        s_spl = CubicSpline(Ts, scales, bc_type="not-a-knot")
        a_spl = CubicSpline(Ts, ampls, bc_type="not-a-knot")

        def lin(spl_, t):
            if t < Ts[0]:
                return float(spl_(Ts[0]) + spl_(Ts[0], 1) * (t - Ts[0]))
            if t > Ts[-1]:
                return float(spl_(Ts[-1]) + spl_(Ts[-1], 1) * (t - Ts[-1]))
            return float(spl_(t))
        for t in (1.7, 4.6, -0.5, 5.4):
            s = 1.0 / lin(s_spl, t)
            a = lin(a_spl, t) if cols == 3 else 0.6
            # Phi_s = a s Phi(s x): the transform with ampl a*s, sigma s
            ref = A.transformed(mp_ref, x, ampl=a * s, sigma=s)
            p, F, D = [q.cpu().numpy() for q in sc.evalDeriv(xt, t=t)]
            rho = sc.density(xt, t=t).cpu().numpy()
            assert np.max(np.abs(p - ref[0]) / np.abs(ref[0])) < 1e-12, (cols, t)
            assert np.max(np.abs(F + ref[1])) < 1e-11 * np.abs(ref[1]).max(), (cols, t)
            R = np.hypot(x[:, 0], x[:, 1])
            ok = (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))
            assert np.max(np.abs(D[ok] + ref[2][ok])) < 1e-10 * np.abs(ref[2][ok]).max(), (cols, t)
            assert np.max(np.abs(rho - ref[3])) < 1e-10 * np.abs(ref[3]).max(), (cols, t)
