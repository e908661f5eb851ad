"""GPU parity of the gfx950 CylSpline kernel against the oracle
(oracle/cylspline_oracle.py) on the reference's bar fixture, inside and
outside the grid, and as a leaf of a fused composite.
Tolerances: phi max|d|/|phi| <= 1e-12; force max|d|/max|F| <= 1e-11;
deriv max|d|/max|D| <= 1e-10 inside the grid; outside, where the oracle's
Hessian is a central difference of a complex-step gradient, 1e-6."""
import numpy as np
import pytest
import torch

from conftest import FIXTURES

pytestmark = pytest.mark.gpu
FILE = FIXTURES / "600.bar.none_8.coef_cylsp_DR"


@pytest.fixture(scope="module")
def setup():
    from agama_gpu_amd import CylSplinePotentialGPU, read_coefs
    from oracle.cylspline_oracle import OracleCylSpline
    c = read_coefs(FILE)
    return CylSplinePotentialGPU(c), OracleCylSpline(c)


def points(n, seed):
    rng = np.random.default_rng(seed)
    r = np.exp(rng.uniform(np.log(0.05), np.log(2500.0), n))
    ct = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    x = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)
    # the reference's smoke points (diag_verify_inner.py:24-31), z axis, plane
    x[:8] = [[1.0, 0.0, 0.0], [2.0, 0.5, 0.5], [5.0, 1.0, 1.0], [8.0, -0.3, 2.0],
             [0.5, 0.0, -0.5], [12.0, 0.0, 0.0], [0.0, 0.0, 3.0], [3.0, 4.0, 0.0]]
    return x


@pytest.fixture(params=[0, 1], ids=["deferred", "per_lane"])
def cyl_path(request):
    """Both CylSpline kernels: points outside the grid deferred to full
    wavefronts (default), and one point per lane."""
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_CYL_PATH, request.param)
    yield request.param
    _lib.set_option(_lib.AGX_OPT_CYL_PATH, 0)


@pytest.mark.parametrize("n", [20000, 1, 255, 257, 70001])
def test_cylspline_parity(setup, cyl_path, n):
    gpu, orc = setup
    x = points(max(n, 8), n)[:n]
    print(f"cylspline N={n} path={cyl_path}", flush=True)
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(torch.as_tensor(x, device="cuda"))]
    P, g, H, outer = orc.eval(x)
    if n >= 255:
        assert 0.05 < outer.mean() < 0.95
    assert np.max(np.abs(p - P) / np.abs(P)) < 1e-12
    assert np.max(np.abs(F + g)) <= 1e-11 * np.abs(g).max()
    ins = ~outer
    if ins.any():
        assert np.max(np.abs(D[ins] + H[ins])) <= 1e-10 * np.abs(H[ins]).max()
    if outer.any():
        assert np.max(np.abs(D[outer] + H[outer])) <= 1e-6 * np.abs(H[outer]).max()
    xt = torch.as_tensor(x, device="cuda")
    rho = gpu.density(xt).cpu().numpy()
    tr = (H[:, 0] + H[:, 1] + H[:, 2]) / (4 * np.pi * 4.30091727067736e-06)
    if ins.any():
        assert np.max(np.abs(rho[ins] - tr[ins])) <= 1e-9 * np.abs(tr[ins]).max()
    assert np.all(rho[outer] == 0.0)
    # pot-only and pot+force kernels
    p1 = gpu.potential(xt).cpu().numpy()
    p2, F2 = [t.cpu().numpy() for t in gpu.eval(xt, pot=True, acc=True)]
    assert np.max(np.abs(p1 - P) / np.abs(P)) < 1e-12 and np.max(np.abs(p2 - P) / np.abs(P)) < 1e-12
    assert np.max(np.abs(F2 + g)) <= 1e-11 * np.abs(g).max()


def test_cylspline_in_fused_composite(setup):
    """Config-1 shape: Multipole + CylSpline + NFW + Plummer, one launch."""
    from agama_gpu_amd import (CompositePotentialGPU, MultipolePotentialGPU, PotentialGPU,
                               read_coefs)
    from oracle import analytic_oracle as A
    from oracle.oracle import OracleMultipole
    gpu, orc = setup
    mc = read_coefs(FIXTURES / "600.dark.none_8.coef_mul_DR")
    comp = CompositePotentialGPU([MultipolePotentialGPU(mc), gpu,
                                  PotentialGPU(type="NFW", mass=1e12, scaleRadius=20.0),
                                  PotentialGPU(type="Plummer", mass=1e10, scaleRadius=0.5)])
    x = points(20000, 1)
    p, F, D = [t.cpu().numpy() for t in comp.evalDeriv(torch.as_tensor(x, device="cuda"))]
    pm, Fm, Dm = OracleMultipole.from_coefs(mc).eval_deriv(x)
    Pc, gc, Hc, outer = orc.eval(x)
    Pn, gn, Hn, _ = A.nfw(x, 1e12, 20.0)
    Pp, gp, Hp, _ = A.plummer(x, 1e10, 0.5)
    P = pm + Pc + Pn + Pp
    Fr = Fm - gc - gn - gp
    Dr = Dm - Hc - Hn - Hp
    assert np.max(np.abs(p - P) / np.abs(P)) < 1e-12
    assert np.max(np.abs(F - Fr)) <= 1e-11 * np.abs(Fr).max()
    R = np.hypot(x[:, 0], x[:, 1])
    ok = ((R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))) & ~outer
    assert np.max(np.abs(D[ok] - Dr[ok])) <= 1e-10 * np.abs(Dr[ok]).max()


def test_cylspline_grid_beyond_lds():
    """A table whose grids (200 + 320 nodes) exceed the 512 nodes the
    deferred kernels copy into LDS: the standalone evaluation takes the
    one-point-per-lane kernel and the composite evaluates the leaf point by
    point; both equal the oracle (the exterior field of exact_fields on a
    finer grid)."""
    import math
    from agama_gpu_amd import CompositePotentialGPU, CylSplinePotentialGPU, PotentialGPU
    from agama_gpu_amd.coefs import CylSplineCoefs
    from exact_fields import exterior_harmonics
    from oracle import analytic_oracle as A
    from oracle.cylspline_oracle import OracleCylSpline
    R = np.concatenate([[0.0], 60.0 * np.sinh(np.linspace(0.05, 2.0, 199)) / math.sinh(2.0)])
    R[-1] = 60.0
    zp = 80.0 * np.sinh(np.linspace(0.02, 2.0, 160)) / math.sinh(2.0)
    zp[-1] = 80.0
    z = np.concatenate([-zp[::-1], zp])
    RR, ZZ = np.meshgrid(R, z, indexing="ij")
    s = np.maximum(1.0, 2.0 / np.hypot(RR, ZZ))
    c = CylSplineCoefs(R, z, {m: np.asarray(t, float)
                              for m, t in exterior_harmonics(RR * s, ZZ * s).items()}, 2)
    gpu, orc = CylSplinePotentialGPU(c), OracleCylSpline(c)
    x = points(20000, 7) * 0.15
    P, g, H, outer = orc.eval(x)
    assert 0.05 < outer.mean() < 0.95
    print("cylspline 200x320 grid, N=20000", flush=True)
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(torch.as_tensor(x, device="cuda"))]
    assert np.max(np.abs(p - P) / np.abs(P)) < 1e-12
    assert np.max(np.abs(F + g)) <= 1e-11 * np.abs(g).max()
    ins = ~outer
    assert np.max(np.abs(D[ins] + H[ins])) <= 1e-10 * np.abs(H[ins]).max()
    comp = CompositePotentialGPU([gpu, PotentialGPU(type="Plummer", mass=1.0 / A.G,
                                                    scaleRadius=0.5)])
    pc, Fc = [t.cpu().numpy() for t in comp.eval(torch.as_tensor(x, device="cuda"), pot=True,
                                                   acc=True)]
    Pp, gp, _, _ = A.plummer(x, 1.0 / A.G, 0.5)
    assert np.max(np.abs(pc - (P + Pp)) / np.abs(P + Pp)) < 1e-12
    assert np.max(np.abs(Fc + g + gp)) <= 1e-11 * np.abs(g + gp).max()


def test_cylspline_outer_on_axis(setup):
    """Points on the z axis (and within sin(theta) < 1e-10 of it) beyond the
    grid box, where the outer multipole's Hessian is its regular limit: equal
    to the oracle's (a central difference of its complex-step gradient), at
    both poles; the chain rule's pole branch (d2Y/dtheta2 = 0, no |m| = 1, 2
    limits) was off there."""
    gpu, orc = setup
    z = np.array([600.0, -600.0, 2000.0, -2500.0])
    x = np.concatenate([np.stack([0 * z, 0 * z, z], axis=1),
                        np.stack([3e-12 * np.abs(z), 0 * z, z], axis=1)])
    print("cylspline outer on axis", flush=True)
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(torch.as_tensor(x, device="cuda"))]
    P, g, H, outer = orc.eval(x)
    assert outer.all()
    assert np.max(np.abs(p - P) / np.abs(P)) < 1e-12
    assert np.max(np.abs(F + g)) <= 1e-11 * np.abs(g).max()
    for i in range(len(x)):
        assert np.max(np.abs(D[i] + H[i])) <= 1e-6 * np.abs(H[i]).max(), (x[i], D[i], -H[i])


@pytest.mark.parametrize("table", ["bar", "m02_log", "m02_linear"])
def test_cylspline_inner_on_axis(setup, cyl_path, table):
    """Points exactly on the z axis inside the grid (R = 0, and R = 1e-31
    where 1/R is still finite) mixed into random points: the Hessian is the
    regular limit (the m = 0 and m = +-2 terms' lR^2 parts; the other
    harmonics' average over directions), equal to the oracle's, which the
    oracle tests pin by differences across the axis; the density equals the
    Hessian's trace / 4 pi G; the same as a composite leaf."""
    from agama_gpu_amd import CompositePotentialGPU, CylSplinePotentialGPU, PotentialGPU
    from oracle import analytic_oracle as A
    from oracle.cylspline_oracle import OracleCylSpline
    if table == "bar":
        gpu, orc = setup
    else:
        from exact_fields import axis_tables
        c = axis_tables()[table[4:]]
        gpu, orc = CylSplinePotentialGPU(c), OracleCylSpline(c)
    rng = np.random.default_rng(11)
    n = 4096
    zmax = orc.Rscale * np.sinh(orc.lz[-1])
    z = rng.uniform(-0.9, 0.9, n) * zmax
    ph = rng.uniform(0, 2 * np.pi, n)
    Rr = orc.Rscale * np.sinh(rng.uniform(orc.lR[0], orc.lR[-1], n))
    R = np.where(np.arange(n) % 3 == 0, 0.0, np.where(np.arange(n) % 3 == 1, 1e-31, Rr))
    x = np.stack([R * np.cos(ph), R * np.sin(ph), z], axis=1)
    print(f"cylspline inner on axis table={table} path={cyl_path}", flush=True)
    P, g, H, outer = orc.eval(x)
    assert not outer.any()
    xt = torch.as_tensor(x, device="cuda")
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(xt)]
    assert np.max(np.abs(p - P) / np.abs(P)) < 1e-12
    assert np.max(np.abs(F + g)) <= 1e-11 * np.abs(g).max()
    axis = R < 1e-30
    for sel in (axis, ~axis):
        assert np.max(np.abs(D[sel] + H[sel])) <= 1e-10 * np.abs(H[sel]).max()
    rho = gpu.density(xt).cpu().numpy()
    tr = (H[:, 0] + H[:, 1] + H[:, 2]) / (4 * np.pi * A.G)
    assert np.max(np.abs(rho - tr)) <= 1e-9 * np.abs(tr).max()
    comp = CompositePotentialGPU([gpu, PotentialGPU(type="Plummer", mass=1.0 / A.G,
                                                    scaleRadius=0.5)])
    pc, Fc, Dc = [t.cpu().numpy() for t in comp.evalDeriv(xt)]
    Pp, gp, Hp, _ = A.plummer(x, 1.0 / A.G, 0.5)
    assert np.max(np.abs(Dc[axis] + H[axis] + Hp[axis])) <= 1e-10 * np.abs(H[axis] + Hp[axis]).max()
    assert np.max(np.abs(Fc + g + gp)) <= 1e-11 * np.abs(g + gp).max()
