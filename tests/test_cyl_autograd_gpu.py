"""Gradients of the CylSpline adjoint kernel (agx_cylspline_adjoint) on gfx950,
on the reference's bar fixture (17 harmonics, 25 x 49 grid, log-scaled m=0)
and, where marked, the Milky-Way disk + bar table (7 harmonics, a 32 x 32
half-space grid mirrored to z < 0: the mirroring of the rebuild).

* forward of the differentiable path = CylSplinePotentialGPU (1e-12 / 1e-11).
* dL/dxyz is a Hessian-vector product: against the forward kernel's own
  evalDeriv, gx = -w_phi F + D w_F (D = -Hessian), to 1e-10 of the largest term.
* dL/d(nodes), dL/d(W) of the kernel alone: a directional derivative along a
  random perturbation of every node quantity (f, f_lR, f_lz, f_lRlz of every
  harmonic) and every outer coefficient, against a Richardson-extrapolated
  FP64 central difference (step 1e-4 of the perturbation), to 1e-7.
* dL/d(table) through the torch rebuild: the same against perturbations of
  every raw Phi_m(R_i, z_j), to 1e-6.
Regions: inside the grid, outside it (the outer multipole), and both; the
z axis and the origin are included.  Upstream weights are seeded normals."""
import numpy as np
import pytest
import torch

from conftest import FIXTURES

pytestmark = pytest.mark.gpu

DEV = "cuda"
FILE = FIXTURES / "600.bar.none_8.coef_cylsp_DR"


def _coefs(name="bar"):
    from agama_gpu_amd.coefs import read_coefs
    from agama_gpu_amd.workloads import mw_disk_bar_coefs
    return read_coefs(FILE) if name == "bar" else mw_disk_bar_coefs()


def _points(region, n=4096, seed=0, ext=500.0):
    """The grid spans R <= ext, |z| <= ext (500 for the bar, 200 for the disk)."""
    rng = np.random.default_rng(seed)
    lo, hi = {"grid": (0.05, 0.9 * ext), "outer": (1.44 * ext, 10.0 * ext),
              "all": (0.05, 6.0 * ext)}[region]
    r = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    ct = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    x = np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)
    if region == "grid":
        x[:3] = [[0.0, 0.0, 0.0], [0.0, 0.0, 3.0], [2.0, 0.5, -0.5]]
    elif region == "outer":
        x[:2] = [[0.0, 0.0, 1.6 * ext], [1.8 * ext, 0.0, 0.0]]
    return x


EXT = {"bar": 500.0, "disk": 200.0}


def _weights(n, seed):
    rng = np.random.default_rng(seed)
    return (torch.as_tensor(rng.normal(size=n), device=DEV),
            torch.as_tensor(rng.normal(size=(n, 3)), device=DEV))


def _loss(out, wphi, wF):
    return (wphi * out[0]).sum() + (wF * out[1]).sum()


@pytest.mark.parametrize("name", ["bar", "disk"])
@pytest.mark.parametrize("region", ["grid", "outer", "all"])
def test_forward_and_xyz_gradient(region, name):
    from agama_gpu_amd import CylSplinePotentialGPU
    from agama_gpu_amd.autograd import FittableCylSpline
    c = _coefs(name)
    x = _points(region, ext=EXT[name])
    print(f"case {name} {region} n={len(x)}", flush=True)
    m = FittableCylSpline(c, device=DEV)
    ref = CylSplinePotentialGPU(c, device=DEV)
    wphi, wF = _weights(len(x), 5)
    xyz = torch.as_tensor(x, device=DEV).requires_grad_(True)
    phi, F = m(xyz)
    (gx,) = torch.autograd.grad(_loss((phi, F), wphi, wF), xyz)
    p0, F0, D = ref.evalDeriv(x)
    assert torch.max(torch.abs(phi - p0) / torch.abs(p0)).item() < 1e-12
    assert (torch.max(torch.abs(F - F0)) / torch.max(torch.abs(F0))).item() < 1e-11
    Dm = torch.stack([torch.stack([D[:, 0], D[:, 3], D[:, 5]], 1),
                      torch.stack([D[:, 3], D[:, 1], D[:, 4]], 1),
                      torch.stack([D[:, 5], D[:, 4], D[:, 2]], 1)], 1)
    want = -wphi[:, None] * F0 + torch.einsum("nij,nj->ni", Dm, wF)
    scale = torch.max(torch.abs(wphi[:, None] * F0)) + torch.max(torch.abs(Dm)) * 3.0
    err = (torch.max(torch.abs(gx - want)) / scale).item()
    assert err < 1e-10, err


def _richardson(central, h=1e-6):
    return (4.0 * central(0.5 * h) - central(h)) / 3.0


@pytest.mark.parametrize("region", ["grid", "outer", "all"])
def test_node_and_outer_gradient_directional(region):
    """The adjoint kernel alone: nodes and W as independent leaves."""
    from agama_gpu_amd.autograd import FittableCylSpline, _CylSplineEval
    c = _coefs()
    x = _points(region, seed=1)
    print(f"case {region} n={len(x)}", flush=True)
    m = FittableCylSpline(c, device=DEV)
    with torch.no_grad():
        nodes0, W0 = m.tables()
    nodes0, W0 = nodes0.contiguous(), W0.contiguous()
    xyz = torch.as_tensor(x, device=DEV)
    wphi, wF = _weights(len(x), 11)
    nodes = nodes0.clone().requires_grad_(True)
    W = W0.clone().requires_grad_(True)
    gN, gW = torch.autograd.grad(_loss(_CylSplineEval.apply(xyz, nodes, W, m), wphi, wF),
                                 [nodes, W])
    assert torch.isfinite(gN).all() and torch.isfinite(gW).all()
    rng = np.random.default_rng(12)
    # per-quantity scale so that slopes and values move alike
    qs = nodes0.abs().amax(dim=(1, 2), keepdim=True)
    vN = qs * torch.as_tensor(rng.normal(size=gN.shape), device=DEV)
    vW = (W0.abs() + 1e-3 * W0.abs().max()) * torch.as_tensor(rng.normal(size=gW.shape), device=DEV)
    ad = ((gN * vN).sum() + (gW * vW).sum()).item()

    def central(eps):
        vals = []
        with torch.no_grad():
            for sgn in (1.0, -1.0):
                out = _CylSplineEval.apply(xyz, (nodes0 + sgn * eps * vN).contiguous(),
                                           (W0 + sgn * eps * vW).contiguous(), m)
                vals.append(_loss(out, wphi, wF).item())
        return (vals[0] - vals[1]) / (2 * eps)

    # The kernel's output is linear in the node values and W (the log
    # scaling of m = 0 aside), so a step of 1e-4 costs nothing in truncation
    # after the Richardson step, while at 1e-6 the round-off of the loss sum
    # (terms much larger than the loss) reaches 1e-7 of the derivative in the
    # outer region.
    fd = _richardson(central, 1e-4)
    assert abs(ad - fd) <= 1e-7 * abs(fd), (ad, fd)
    if region == "grid":
        assert gW.abs().max().item() == 0.0
    if region == "outer":
        assert gN.abs().max().item() == 0.0


@pytest.mark.parametrize("name", ["bar", "disk"])
@pytest.mark.parametrize("region", ["grid", "outer", "all"])
def test_table_gradient_directional(region, name):
    """Through the rebuild: every raw Phi_m(R_i, z_j) entry perturbed."""
    from agama_gpu_amd.autograd import FittableCylSpline
    c = _coefs(name)
    x = _points(region, seed=2, ext=EXT[name])
    print(f"case {name} {region} n={len(x)}", flush=True)
    m = FittableCylSpline(c, device=DEV)
    xyz = torch.as_tensor(x, device=DEV)
    wphi, wF = _weights(len(x), 13)
    (g,) = torch.autograd.grad(_loss(m(xyz), wphi, wF), [m.phi_table])
    assert torch.isfinite(g).all()
    rng = np.random.default_rng(14)
    P0 = m.phi_table.detach().clone()
    # relative per entry, plus 1e-3 of each harmonic's largest entry
    v = (P0.abs() + 1e-3 * P0.abs().amax(dim=(1, 2), keepdim=True)) * \
        torch.as_tensor(rng.normal(size=g.shape), device=DEV)
    ad = (g * v).sum().item()

    def central(eps):
        vals = []
        with torch.no_grad():
            for sgn in (1.0, -1.0):
                m.phi_table.copy_(P0 + sgn * eps * v)
                vals.append(_loss(m(xyz), wphi, wF).item())
            m.phi_table.copy_(P0)
        return (vals[0] - vals[1]) / (2 * eps)

    fd = _richardson(central)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_gradient_repeatable_and_force_only():
    """Two backward passes agree to round-off; a loss of the force alone
    (no dL/dPhi) runs the adjoint with a NULL w_phi."""
    from agama_gpu_amd.autograd import FittableCylSpline
    c = _coefs()
    x = _points("all", seed=3)
    m = FittableCylSpline(c, device=DEV)
    xyz = torch.as_tensor(x, device=DEV)
    wphi, wF = _weights(len(x), 3)
    g1 = torch.autograd.grad(_loss(m(xyz), wphi, wF), m.phi_table)[0]
    g2 = torch.autograd.grad(_loss(m(xyz), wphi, wF), m.phi_table)[0]
    assert (torch.max(torch.abs(g1 - g2)) / torch.max(torch.abs(g1))).item() < 1e-12
    gF = torch.autograd.grad((wF * m(xyz)[1]).sum(), m.phi_table)[0]
    gF0 = torch.autograd.grad(_loss(m(xyz), torch.zeros_like(wphi), wF), m.phi_table)[0]
    assert (torch.max(torch.abs(gF - gF0)) / torch.max(torch.abs(gF0))).item() < 1e-12


def test_gradient_additive_over_batches():
    """Large N (runs of several chunks per wavefront, many cells): the table
    gradient of the whole set equals the sum over three unequal batches, and
    dL/dx matches row by row."""
    from agama_gpu_amd.autograd import FittableCylSpline
    n = 600_000
    x = _points("all", n=n, seed=9)
    print(f"case bar n={n}", flush=True)
    m = FittableCylSpline(_coefs(), device=DEV)
    xyz = torch.as_tensor(x, device=DEV)
    wphi, wF = _weights(n, 10)

    def grads(sl):
        xs = xyz[sl].clone().requires_grad_(True)
        out = m(xs)
        return torch.autograd.grad(_loss(out, wphi[sl], wF[sl]), [m.phi_table, xs])

    g, gx = grads(slice(0, n))
    acc = torch.zeros_like(g)
    gx_parts = []
    for sl in (slice(0, 100), slice(100, 250_000), slice(250_000, n)):
        a, gxs = grads(sl)
        acc += a
        gx_parts.append(gxs)
    assert (torch.max(torch.abs(g - acc)) / torch.max(torch.abs(g))).item() < 1e-11
    assert torch.equal(gx, torch.cat(gx_parts))
