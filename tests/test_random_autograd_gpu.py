"""Multipole adjoint on random table structures (tests/random_tables.py):
sin-only odd-m groups, sparse subsets, lmax up to 12, a sign-changing
monopole without log-scaling.  Same checks as the fixture tests
(tests/test_autograd_gpu.py): dL/dx equals the Hessian-vector product formed
from the forward kernel's evalDeriv (1e-10 of the largest term, rows near the
z axis excluded), and the table gradient along a random direction of every
raw entry equals a Richardson-extrapolated central difference (1e-6).
"""
import numpy as np
import pytest
import torch

from random_tables import random_multipole, sweep_points
from test_autograd_gpu import _loss

pytestmark = pytest.mark.gpu

DEV = "cuda"
SEEDS = [3, 4, 5, 12, 17]


@pytest.mark.parametrize("seed", SEEDS)
def test_random_table_xyz_gradient(seed):
    from agama_gpu_amd import MultipolePotentialGPU
    from agama_gpu_amd.autograd import FittableMultipole
    c, d = random_multipole(seed)
    x = sweep_points(d, 4096, seed)
    print(f"case seed={seed} lmax={d['lmax']} pattern={d['pattern']} signed={d['signed']}",
          flush=True)
    m = FittableMultipole(c, device=DEV)
    ref = MultipolePotentialGPU(c, device=DEV)
    rng = np.random.default_rng(seed)
    wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
    xyz = torch.as_tensor(x, device=DEV).requires_grad_(True)
    phi, F = m(xyz)
    (gx,) = torch.autograd.grad((wphi * phi).sum() + (wF * F).sum(), xyz)
    p0, F0, D = ref.evalDeriv(x)
    assert (torch.max(torch.abs(phi - p0)) / torch.max(torch.abs(p0))).item() < 1e-12
    assert (torch.max(torch.abs(F - F0)) / torch.max(torch.abs(F0))).item() < 1e-11
    Dm = torch.stack([torch.stack([D[:, 0], D[:, 3], D[:, 5]], 1),
                      torch.stack([D[:, 3], D[:, 1], D[:, 4]], 1),
                      torch.stack([D[:, 5], D[:, 4], D[:, 2]], 1)], 1)
    want = -wphi[:, None] * F0 + torch.einsum("nij,nj->ni", Dm, wF)
    R = torch.sqrt(xyz[:, 0] ** 2 + xyz[:, 1] ** 2).detach()
    r = torch.sqrt(R * R + xyz[:, 2].detach() ** 2)
    ok = ~((R > 0) & (R < 1e-6 * r))
    scale = torch.max(torch.abs(wphi[:, None] * F0)) + torch.max(torch.abs(Dm)) * 3.0
    assert (torch.max(torch.abs(gx - want)[ok]) / scale).item() < 1e-10, d


@pytest.mark.parametrize("seed", SEEDS)
def test_random_table_gradient_directional(seed):
    from agama_gpu_amd.autograd import FittableMultipole
    c, d = random_multipole(seed)
    x = sweep_points(d, 2048, seed + 1)
    print(f"case seed={seed} lmax={d['lmax']} pattern={d['pattern']} signed={d['signed']}",
          flush=True)
    m = FittableMultipole(c, device=DEV)
    rng = np.random.default_rng(seed + 2)
    xyz = torch.as_tensor(x, device=DEV)
    wphi = torch.as_tensor(rng.normal(size=len(x)), device=DEV)
    wF = torch.as_tensor(rng.normal(size=(len(x), 3)), device=DEV)
    gP, gD = torch.autograd.grad(_loss(m, xyz, wphi, wF), [m.phi_table, m.dphi_table])
    assert torch.isfinite(gP).all() and torch.isfinite(gD).all()
    P0, D0 = m.phi_table.detach().clone(), m.dphi_table.detach().clone()
    vP = (P0.abs() + 1e-3 * P0.abs().max()) * torch.as_tensor(rng.normal(size=gP.shape), device=DEV)
    vD = (D0.abs() + 1e-3 * D0.abs().max()) * torch.as_tensor(rng.normal(size=gD.shape), device=DEV)
    ad = ((gP * vP).sum() + (gD * vD).sum()).item()

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

    fd = (4.0 * central(5e-7) - central(1e-6)) / 3.0
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd, d)
