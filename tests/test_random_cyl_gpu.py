"""GPU parity on random CylSpline table structures (tests/random_tables.py):
twelve seeded tables (mmax 0..6; even, cos-only, all and random subsets of
harmonics; half grids mirrored by the build and full z-asymmetric grids) at
points inside and outside the grid box, near and on the z axis and in the
midplane, through both CylSpline kernels (outside points deferred to full
wavefronts, and one point per lane) and as a shifted, scaled composite leaf.
Tolerances of the fixture tests: phi 1e-12 relative (to max|phi|: the m != 0
harmonics may take the potential through zero), force 1e-11 of max|F|,
deriv 1e-10 of max|D| inside the box (1e-6 outside, where the oracle's
Hessian is a central difference of a complex-step gradient; rows with
0 < sin(theta) < 1e-6 excluded as in the Multipole tests), density equal
to the oracle's Hessian trace / (4 pi G) inside, 0 outside.  Inside the box
the rows with 0 < R < 1e-6 Rscale are held to 1e-7 only: there the Hessian of
the harmonics other than m = 0, +-2 depends on the direction of approach
(the splines have zero lR slope on the axis), grows as R -> 0 and is the
difference of terms of size Phi_m / R^2, so two FP64 evaluations in
different orders differ by ~1e-9 of max|D| (measured 2e-9; beyond 1e-6
Rscale 2e-13).
"""
import numpy as np
import pytest
import torch

from random_tables import cyl_sweep_points, random_cylspline

pytestmark = pytest.mark.gpu

SEEDS = list(range(12))
N = 20_000
FOURPIG = 4 * np.pi * 4.30091727067736e-06


@pytest.fixture(scope="module")
def tables():
    from oracle.cylspline_oracle import OracleCylSpline
    out = {}
    for s in SEEDS:
        c, d = random_cylspline(s)
        x = cyl_sweep_points(d, N, s)
        o = OracleCylSpline(c)
        out[s] = (c, d, x, o, o.eval(x))
    return out


@pytest.fixture(params=[0, 1], ids=["deferred", "per_lane"])
def cyl_path(request):
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_CYL_PATH, request.param)
    yield request.param
    _lib.set_option(_lib.AGX_OPT_CYL_PATH, 0)


def CylSplineRscale(c):
    from agama_gpu_amd.cylspline_prep import build_cylspline_data
    return build_cylspline_data(c)["Rscale"]


def regular(x):
    R = np.hypot(x[:, 0], x[:, 1])
    return (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))


@pytest.mark.parametrize("seed", SEEDS)
def test_random_cylspline_paths(tables, seed, cyl_path):
    from agama_gpu_amd import CylSplinePotentialGPU
    c, d, x, _, (P, g, H, outer) = tables[seed]
    print(f"case seed={seed} mmax={d['mmax']} pattern={d['pattern']} half={d['half']} "
          f"grid={d['nR']}x{d['nz']} path={cyl_path} n={len(x)}", flush=True)
    gpu = CylSplinePotentialGPU(c)
    xt = torch.as_tensor(x, device="cuda")
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(xt)]
    assert np.max(np.abs(p - P)) < 1e-12 * np.abs(P).max(), d
    assert np.max(np.abs(F + g)) <= 1e-11 * np.abs(g).max(), d
    assert np.all(np.isfinite(D)), d
    R = np.hypot(x[:, 0], x[:, 1])
    band = (R > 0) & (R < 1e-6 * gpu._data["Rscale"]) & ~outer
    ins, out = ~outer & ~band, regular(x) & outer
    assert np.max(np.abs(D[ins] + H[ins])) <= 1e-10 * np.abs(H[ins]).max(), d
    if band.any():
        assert np.max(np.abs(D[band] + H[band])) <= 1e-7 * np.abs(H[~outer]).max(), d
    assert np.max(np.abs(D[out] + H[out])) <= 1e-6 * np.abs(H[out]).max(), d
    rho = gpu.density(xt).cpu().numpy()
    tr = (H[:, 0] + H[:, 1] + H[:, 2]) / FOURPIG
    assert np.max(np.abs(rho[ins] - tr[ins])) <= 1e-9 * np.abs(tr[ins]).max(), d
    assert np.all(rho[outer] == 0.0), d
    p2, F2 = [t.cpu().numpy() for t in gpu.eval(xt, pot=True, acc=True)]
    assert np.max(np.abs(p2 - P)) < 1e-12 * np.abs(P).max(), d
    assert np.max(np.abs(F2 + g)) <= 1e-11 * np.abs(g).max(), d


@pytest.mark.parametrize("seed", SEEDS[::2])
def test_random_cylspline_as_composite_leaf(tables, seed):
    """0.7 * s * Phi(s (x - c)) with s = 1/2 inside a composite with a
    Plummer sphere, in one fused launch."""
    from agama_gpu_amd import (CompositePotentialGPU, CylSplinePotentialGPU, PotentialGPU,
                               ScaledPotentialGPU, ShiftedPotentialGPU)
    from oracle import analytic_oracle as A
    c, d, x, o, _ = tables[seed]
    print(f"case seed={seed} composite leaf", flush=True)
    ctr = np.array([0.01, -0.02, 0.015]) * d["R_max"]
    comp = CompositePotentialGPU([
        ShiftedPotentialGPU(ScaledPotentialGPU(CylSplinePotentialGPU(c), 2.0, ampl=0.7), ctr),
        PotentialGPU(type="Plummer", mass=1e9, scaleRadius=0.3 * d["R_max"])])
    p, F, D = [t.cpu().numpy() for t in comp.evalDeriv(torch.as_tensor(x, device="cuda"))]
    s = 0.5
    Pc, gc, Hc, outer = o.eval(s * (x - ctr))
    Pp, gp, Hp, _ = A.plummer(x, 1e9, 0.3 * d["R_max"])
    Pr = 0.7 * s * Pc + Pp
    Fr = -(0.7 * s * s * gc + gp)
    Dr = -(0.7 * s ** 3 * Hc + Hp)
    assert np.max(np.abs(p - Pr)) < 1e-12 * np.abs(Pr).max(), d
    assert np.max(np.abs(F - Fr)) <= 1e-11 * np.abs(Fr).max(), d
    xs = s * (x - ctr)
    band = np.hypot(xs[:, 0], xs[:, 1]) < 1e-6 * CylSplineRscale(c)
    ok = regular(x - ctr) & ~outer & ~band
    assert np.max(np.abs(D[ok] - Dr[ok])) <= 1e-10 * np.abs(Dr[ok]).max(), d
