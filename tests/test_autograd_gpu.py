"""Gradients of the Multipole adjoint kernel (agx_multipole_adjoint) on gfx950.

* dL/dxyz is a Hessian-vector product: checked against the forward kernel's
  own evalDeriv output, gx = -w_phi F + D w_F (D = -Hessian), to 1e-10
  of the largest term (the tolerance of the deriv parity tests).
* dL/d(table) through the torch rebuild and the adjoint scatter: a
  directional derivative along a random perturbation of every
  Phi_lm(r_i), dPhi_lm/dr(r_i) entry (zero harmonics included) against an
  FP64 central difference (step 1e-6 relative), to 1e-6.
* dL/d(knots) of a spline-shifted Multipole: central differences per knot.
Points cover the interior, both extrapolation regions, the z axis and the
origin.  The upstream weights are seeded random normals.
"""
import numpy as np
import pytest
import torch

from conftest import FIXTURES

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _coefs(name):
    from agama_gpu_amd.coefs import read_coefs
    from agama_gpu_amd.workloads import halo_lmax8_coefs
    if name == "halo_lmax8":
        return halo_lmax8_coefs()
    return read_coefs(FIXTURES / {"fire_dense": "600.dark.none_8.coef_mul_DR",
                                  "lmc": "100.LMC.none_8.coef_mult"}[name])


def _points(coefs, n=4096, seed=0):
    from agama_gpu_amd.workloads import hernquist_points
    rng = np.random.default_rng(seed + 17)
    x = hernquist_points(n, seed)
    r = np.asarray(coefs.R_grid)
    u = rng.normal(size=(64, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x[-64:-32] = u[:32] * (r[0] * rng.uniform(0.05, 0.9, 32))[:, None]     # inner law
    x[-32:] = u[32:] * (r[-1] * rng.uniform(1.1, 20.0, 32))[:, None]       # outer law
    x[:4] = [[0, 0, 0], [0, 0, 3.0], [0, 0, -3.0], [5.0, -2.0, 0]]
    return x


def _loss(model, xyz, wphi, wF, t=None):
    out = model(xyz) if t is None else model(xyz, t)
    return (wphi * out[0]).sum() + (wF * out[1]).sum()


@pytest.mark.parametrize("name", ["fire_dense", "lmc", "halo_lmax8"])
def test_xyz_gradient_is_hessian_vector_product(name):
    from agama_gpu_amd import MultipolePotentialGPU
    from agama_gpu_amd.autograd import FittableMultipole
    c = _coefs(name)
    x = _points(c)
    print(f"case {name} n={len(x)}", flush=True)
    m = FittableMultipole(c, device=DEV)
    ref = MultipolePotentialGPU(c, device=DEV)
    rng = np.random.default_rng(5)
    wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
    xyz = torch.as_tensor(x, device=DEV).requires_grad_(True)
    phi, F = m(xyz)
    (gx,) = torch.autograd.grad((wphi * phi).sum() + (wF * F).sum(), xyz)
    p0, F0, D = ref.evalDeriv(x)
    # forward of the differentiable path = plain evaluation
    assert torch.max(torch.abs(phi - p0) / torch.abs(p0)).item() < 1e-12
    assert (torch.max(torch.abs(F - F0)) / torch.max(torch.abs(F0))).item() < 1e-11
    Dm = torch.stack([torch.stack([D[:, 0], D[:, 3], D[:, 5]], 1),
                      torch.stack([D[:, 3], D[:, 1], D[:, 4]], 1),
                      torch.stack([D[:, 5], D[:, 4], D[:, 2]], 1)], 1)
    want = -wphi[:, None] * F0 + torch.einsum("nij,nj->ni", Dm, wF)
    R = torch.sqrt(xyz[:, 0] ** 2 + xyz[:, 1] ** 2).detach()
    r = torch.sqrt(R * R + xyz[:, 2].detach() ** 2)
    ok = ~((R > 0) & (R < 1e-6 * r))          # ill-conditioned near-axis Hessians
    scale = torch.max(torch.abs(wphi[:, None] * F0)) + torch.max(torch.abs(Dm)) * 3.0
    err = (torch.max(torch.abs(gx - want)[ok]) / scale).item()
    assert err < 1e-10, err


def _region_points(coefs, region, n=2048, seed=1):
    """Points only inside the inner law, the outer law, or the grid."""
    rng = np.random.default_rng(seed)
    r = np.asarray(coefs.R_grid)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    lo, hi = {"inner": (r[0] * 0.05, r[0] * 0.9), "outer": (r[-1] * 1.1, r[-1] * 20.0),
              "grid": (r[0] * 1.01, r[-1] * 0.99)}[region]
    return u * np.exp(rng.uniform(np.log(lo), np.log(hi), n))[:, None]


@pytest.mark.parametrize("region", ["all", "inner", "outer", "grid"])
@pytest.mark.parametrize("name", ["fire_dense", "lmc", "halo_lmax8"])
def test_table_gradient_directional(name, region):
    from agama_gpu_amd.autograd import FittableMultipole
    c = _coefs(name)
    x = _points(c, seed=1) if region == "all" else _region_points(c, region)
    print(f"case {name} {region} n={len(x)}", flush=True)
    m = FittableMultipole(c, device=DEV)
    rng = np.random.default_rng(11)
    xyz = torch.as_tensor(x, device=DEV)
    wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
    L = _loss(m, xyz, wphi, wF)
    gP, gD = torch.autograd.grad(L, [m.phi_table, m.dphi_table])
    assert torch.isfinite(gP).all() and torch.isfinite(gD).all()
    # relative perturbation of every entry, plus an absolute one (1e-3 of the
    # table's largest entry) so that harmonics that are zero in the table move too
    P, D = m.phi_table.detach(), m.dphi_table.detach()
    vP = (P.abs() + 1e-3 * P.abs().max()) * torch.as_tensor(rng.normal(size=gP.shape), device=DEV)
    vD = (D.abs() + 1e-3 * D.abs().max()) * torch.as_tensor(rng.normal(size=gD.shape), device=DEV)
    ad = ((gP * vP).sum() + (gD * vD).sum()).item()
    P0, D0 = m.phi_table.detach().clone(), m.dphi_table.detach().clone()

    def central(eps):
        vals = []
        with torch.no_grad():
            for sgn in (1.0, -1.0):
                m.phi_table.copy_(P0 + sgn * eps * vP)
                m.dphi_table.copy_(D0 + sgn * eps * vD)
                vals.append(_loss(m, xyz, wphi, wF).item())
            m.phi_table.copy_(P0)
            m.dphi_table.copy_(D0)
        return (vals[0] - vals[1]) / (2 * eps)

    # Richardson-extrapolated central difference (steps 1e-6, 5e-7): the
    # log-scaled monopole f0 = ln(invPhi0 - 1/Phi0) of a table whose inner
    # law nearly cancels (FIRE: invPhi0 Phi0(r_0) = 1 - 1.3e-4) is so curved
    # that a plain step-1e-6 difference carries a 1e-5 truncation error.
    fd = (4.0 * central(5e-7) - central(1e-6)) / 3.0
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_table_gradient_repeatable():
    """Two backward passes (atomics in a different order) agree to round-off."""
    from agama_gpu_amd.autograd import FittableMultipole
    c = _coefs("fire_dense")
    x = _points(c, seed=2)
    m = FittableMultipole(c, device=DEV)
    xyz = torch.as_tensor(x, device=DEV)
    rng = np.random.default_rng(2)
    wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
    g1 = torch.autograd.grad(_loss(m, xyz, wphi, wF), m.phi_table)[0]
    g2 = torch.autograd.grad(_loss(m, xyz, wphi, wF), m.phi_table)[0]
    assert (torch.max(torch.abs(g1 - g2)) / torch.max(torch.abs(g1))).item() < 1e-12


@pytest.mark.parametrize("t", [0.35, 2.6])
def test_knot_gradient(t):
    from agama_gpu_amd.autograd import FittableMultipole, FittableShiftedMultipole, SplineCenter
    c = _coefs("halo_lmax8")
    x = _points(c, n=2048, seed=3)
    rng = np.random.default_rng(4)
    times = np.linspace(-1.0, 2.0, 7)
    knots = rng.normal(size=(7, 3)) * 5.0
    model = FittableShiftedMultipole(FittableMultipole(c, device=DEV),
                                     SplineCenter(times, knots, device=DEV))
    xyz = torch.as_tensor(x, device=DEV)
    wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
    g = torch.autograd.grad(_loss(model, xyz, wphi, wF, t), model.center.knots)[0].cpu().numpy()
    K0 = model.center.knots.detach().clone()
    eps = 1e-5
    for (i, j) in [(2, 0), (3, 1), (5, 2), (6, 0)]:
        vals = []
        with torch.no_grad():
            for sgn in (1.0, -1.0):
                K = K0.clone()
                K[i, j] += sgn * eps
                model.center.knots.copy_(K)
                vals.append(_loss(model, xyz, wphi, wF, t).item())
            model.center.knots.copy_(K0)
        fd = (vals[0] - vals[1]) / (2 * eps)
        scale = max(abs(fd), np.abs(g).max())
        assert abs(g[i, j] - fd) <= 1e-6 * scale, (i, j, g[i, j], fd)


def test_gradient_additive_over_batches():
    """Large N (runs of several chunks per wavefront, slabs flushed across
    intervals): the table gradient of the whole set equals the sum over
    three unequal batches, and dL/dx matches row by row."""
    from agama_gpu_amd.autograd import FittableMultipole
    c = _coefs("fire_dense")
    n = 600_000
    x = _points(c, n=n, seed=7)
    print(f"case fire_dense n={n}", flush=True)
    m = FittableMultipole(c, device=DEV)
    rng = np.random.default_rng(8)
    xyz = torch.as_tensor(x, device=DEV)
    wphi = torch.as_tensor(rng.normal(size=n), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(n, 3)), device=DEV)

    def grads(sl):
        xs = xyz[sl].clone().requires_grad_(True)
        return torch.autograd.grad(_loss(m, xs, wphi[sl], wF[sl]), [m.phi_table, m.dphi_table, xs])

    gP, gD, gx = grads(slice(0, n))
    parts = [slice(0, 100), slice(100, 250_000), slice(250_000, n)]
    acc = [torch.zeros_like(gP), torch.zeros_like(gD)]
    gx_parts = []
    for sl in parts:
        a, b, gxs = grads(sl)
        acc[0] += a
        acc[1] += b
        gx_parts.append(gxs)
    for full, summed in zip((gP, gD), acc):
        assert (torch.max(torch.abs(full - summed)) / torch.max(torch.abs(full))).item() < 1e-11
    assert torch.equal(gx, torch.cat(gx_parts))


# (inner_mode, inner_reset, outer_mode, log_scaling) forced on a fixture: every
# branch of the device rebuild and its adjoint against the torch rebuild
MODES = [None, ("log", False, "log", 1), ("const", False, "kepler", 1),
         ("s2", True, "kepler", 1), ("s2", False, "log", 0)]


@pytest.mark.parametrize("modes", MODES, ids=["fixture", "log-log", "const-kepler", "s2reset-kepler",
                                              "linear"])
@pytest.mark.parametrize("name", ["fire_dense", "lmc", "halo_lmax8"])
def test_device_build_matches_torch_rebuild(name, modes):
    """agx_multipole_build / _adjoint equal the torch rebuild and its autograd
    (poly and law to 1e-12 of each column's scale, gradients to 1e-11)."""
    from agama_gpu_amd.autograd import FittableMultipole
    c = _coefs(name)
    print(f"case {name} modes={modes}", flush=True)
    m = FittableMultipole(c, device=DEV)
    assert m._device_build
    if modes is not None:
        im, reset, om, ls = modes
        m._d = dict(m._d, inner_mode=im, inner_reset=reset, outer_mode=om, log_scaling=ls)
        m._modes = ({"log": 1, "s2": 2}.get(im, 0), int(reset), 1 if om == "log" else 0)
    poly_d, law_d = m.tables()
    poly_t, law_t = m.tables_torch()
    scale = poly_t.abs().amax(dim=(0, 2), keepdim=True).clamp_min(1e-300)
    assert (torch.max(torch.abs(poly_d - poly_t) / scale)).item() < 1e-12
    assert torch.allclose(law_d, law_t, rtol=1e-13, atol=0.0)
    rng = np.random.default_rng(3)
    Gp = torch.as_tensor(rng.normal(size=poly_t.shape), device=DEV)
    Gl = torch.as_tensor(rng.normal(size=5), device=DEV)
    gd = torch.autograd.grad((poly_d * Gp).sum() + (law_d * Gl).sum(), [m.phi_table, m.dphi_table])
    gt = torch.autograd.grad((poly_t * Gp).sum() + (law_t * Gl).sum(), [m.phi_table, m.dphi_table])
    for a, b in zip(gd, gt):
        assert (torch.max(torch.abs(a - b)) / torch.max(torch.abs(b))).item() < 1e-11


def test_backward_reuses_forward_sort():
    """With the interval-sorted forward (AGX_OPT_MP_PATH 4) the backward
    traverses the forward's sort (AGX_WORK_SORTED) instead of sorting again:
    same gradients as with an LDS-path forward, to round-off."""
    from agama_gpu_amd import _lib
    from agama_gpu_amd.autograd import FittableMultipole
    c = _coefs("halo_lmax8")
    x = _points(c, n=50_000, seed=4)
    m = FittableMultipole(c, device=DEV)
    rng = np.random.default_rng(6)
    xyz = torch.as_tensor(x, device=DEV)
    wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
    out = []
    try:
        for path in (0, 4):
            _lib.set_option(_lib.AGX_OPT_MP_PATH, path)
            xs = xyz.clone().requires_grad_(True)
            out.append(torch.autograd.grad(_loss(m, xs, wphi, wF),
                                           [m.phi_table, m.dphi_table, xs]))
    finally:
        _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)
    for a, b in zip(*out):
        assert (torch.max(torch.abs(a - b)) / torch.max(torch.abs(a))).item() < 1e-12


@pytest.mark.parametrize("n", [2048, 300_000])
def test_fused_shift_matches_explicit_shift(n):
    """The shifted fit evaluates x - c inside the Multipole evaluation and takes
    the centre's gradient from the adjoint kernel's own sum of dL/dx
    (AGX_ADJ_CENTER_GRAD); the same model through an explicit shift (a torch
    subtraction and column sums) gives the same knot, table and position
    gradients."""
    from agama_gpu_amd.autograd import (FittableMultipole, FittableShiftedMultipole, SplineCenter,
                                        _Shift)
    c = _coefs("halo_lmax8")
    x = _points(c, n=n, seed=5)
    print(f"fused shift n={n}", flush=True)
    rng = np.random.default_rng(6)
    times = np.linspace(-1.0, 2.0, 7)
    model = FittableShiftedMultipole(FittableMultipole(c, device=DEV),
                                     SplineCenter(times, rng.normal(size=(7, 3)) * 5.0, device=DEV))
    wphi = torch.as_tensor(rng.normal(size=n), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(n, 3)), device=DEV)
    params = [model.center.knots, model.multipole.phi_table, model.multipole.dphi_table]

    def grads(fn):
        xyz = torch.as_tensor(x, device=DEV).requires_grad_(True)
        phi, F = fn(xyz)
        loss = (wphi * phi).sum() + (wF * F).sum()
        return torch.autograd.grad(loss, params + [xyz])

    fused = grads(lambda xyz: model(xyz, 0.8))
    explicit = grads(lambda xyz: model.multipole(_Shift.apply(xyz, model.center(0.8))))
    for a, b in zip(fused, explicit):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.max(np.abs(a - b)) <= 1e-11 * np.abs(b).max()


def test_adjoint_staged_sums_equal_cross_lane_sums():
    """The two table-gradient reductions of the adjoint (terms staged in LDS
    eight at a time, the default; six cross-lane sums per term,
    AGX_OPT_ADJ_STAGE 0) give the same gradients to round-off, for a
    stride-1 table with sin terms (FIRE) and the stride-2 halo table."""
    from agama_gpu_amd import _lib
    from agama_gpu_amd.autograd import FittableMultipole
    for name in ("fire_dense", "halo_lmax8"):
        c = _coefs(name)
        x = _points(c, n=50_000, seed=9)
        m = FittableMultipole(c, device=DEV)
        rng = np.random.default_rng(9)
        xyz = torch.as_tensor(x, device=DEV)
        wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
        wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
        out = {}
        try:
            for stage in (0, 1):
                _lib.set_option(_lib.AGX_OPT_ADJ_STAGE, stage)
                out[stage] = torch.autograd.grad(_loss(m, xyz, wphi, wF),
                                                 [m.phi_table, m.dphi_table])
        finally:
            _lib.set_option(_lib.AGX_OPT_ADJ_STAGE, 1)
        for a, b in zip(out[0], out[1]):
            assert (torch.max(torch.abs(a - b)) / torch.max(torch.abs(a))).item() < 1e-12, name
