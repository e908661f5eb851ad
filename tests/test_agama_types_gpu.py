"""GPU evaluation of the types the reference builds with Agama's CPU code
(gpu_potential.py:1935-2052), here built by this package (expansion.py) and
evaluated by the Multipole / DiskAnsatz kernels through PotentialGPU(type=...):

* Spheroid (triaxial Dehnen-like and Plummer-like), King and Disk: the GPU
  equals the oracle on the same tables (phi 1e-12, force 1e-11, deriv 1e-10
  of the maximum, the parity bar of the Multipole tests);
* against the mathematics: a spherical Plummer Spheroid gives -GM/sqrt(r^2+b^2)
  and King -GM/r beyond its tidal radius (both within the tables' spline
  interpolation error), and the Disk composite gives the Kuijken-Gilmore
  potential of a double-exponential disk to 2e-5.
"""
import math

import numpy as np
import pytest
import torch

from oracle import analytic_oracle as A

pytestmark = pytest.mark.gpu
DEV = "cuda"
G = A.G


def _pts(n, seed, a=20.0):
    from agama_gpu_amd.workloads import hernquist_points
    return hernquist_points(n, seed, a=a)


def _parity(gpu_out, ref):
    p, F, D = [t.cpu().numpy() for t in gpu_out]
    po, Fo, Do = ref
    assert np.max(np.abs(p - po) / np.abs(po)) < 1e-12
    assert np.max(np.abs(F - Fo)) < 1e-11 * np.abs(Fo).max()
    assert np.max(np.abs(D - Do)) < 1e-10 * np.abs(Do).max()


def test_spheroid_triaxial():
    from agama_gpu_amd import PotentialGPU
    from agama_gpu_amd.expansion import spheroid_coefs
    from oracle.oracle import OracleMultipole
    kw = dict(mass=1e12, scaleRadius=20.0, alpha=1.0, beta=4.0, gamma=1.0, axisRatioY=0.8,
              axisRatioZ=0.6)
    pot = PotentialGPU(type="Spheroid", lmax=8, **kw)
    x = _pts(20000, 31)
    orc = OracleMultipole.from_coefs(spheroid_coefs(lmax=8, **kw))
    R = np.hypot(x[:, 0], x[:, 1])
    ok = (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))
    x = x[ok]
    _parity(pot.evalDeriv(torch.as_tensor(x, device=DEV)), orc.eval_deriv(x))


def test_spheroid_plummer_closed_form():
    from agama_gpu_amd import PotentialGPU
    M, b = 1e10, 0.5
    pot = PotentialGPU(type="Spheroid", mass=M, scaleRadius=b, alpha=2, beta=5, gamma=0)
    x = _pts(20000, 32, a=1.0)
    r = np.linalg.norm(x, axis=1)
    x = x[(r > 5e-4) & (r < 500.0)]          # inside the table's grid
    r = np.linalg.norm(x, axis=1)
    p, F = [t.cpu().numpy() for t in pot.eval(torch.as_tensor(x, device=DEV), pot=True, acc=True)]
    s = np.sqrt(r * r + b * b)
    # within the quintic spline's interpolation error (50 nodes over 13.8
    # e-folds in r; 3e-8 and 5e-7 on the oracle)
    assert np.max(np.abs(p / (-G * M / s) - 1)) < 1e-7
    Fe = -(G * M / s ** 3)[:, None] * x
    assert np.max(np.abs(F - Fe)) < 2e-6 * np.abs(Fe).max()


def test_king():
    from agama_gpu_amd import PotentialGPU
    from agama_gpu_amd.expansion import king_coefs, king_profile
    from oracle.oracle import OracleMultipole
    M, r0, W0 = 1e5, 0.01, 7.0
    pot = PotentialGPU(type="King", mass=M, scaleRadius=r0, W0=W0)
    orc = OracleMultipole.from_coefs(king_coefs(mass=M, scaleRadius=r0, W0=W0))
    _, xi_t = king_profile(W0)
    rt = xi_t * r0
    x = _pts(20000, 33, a=r0)
    R = np.hypot(x[:, 0], x[:, 1])
    x = x[(R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))]
    _parity(pot.evalDeriv(torch.as_tensor(x, device=DEV)), orc.eval_deriv(x))
    # beyond the tidal radius: -GM/r (inside the table's grid and beyond it)
    u = np.random.default_rng(5).normal(size=(4096, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    rr = rt * np.exp(np.random.default_rng(6).uniform(math.log(1.5), math.log(1e3), 4096))
    y = u * rr[:, None]
    p, F = [t.cpu().numpy() for t in pot.eval(torch.as_tensor(y, device=DEV), pot=True, acc=True)]
    # the table samples -GM/r there; within its interpolation error (4e-8 and
    # 5e-7 on the oracle), exact beyond its last node
    assert np.max(np.abs(p / (-G * M / rr) - 1)) < 1e-7
    assert np.max(np.abs(F + (G * M / rr ** 3)[:, None] * y)) < 2e-6 * (G * M / (1.5 * rt) ** 2)


def test_disk():
    from agama_gpu_amd import CompositePotentialGPU, PotentialGPU
    from agama_gpu_amd.expansion import disk_potential_parts
    from oracle.oracle import OracleMultipole
    from test_expansion import _kuijken_gilmore
    S, Rd, h = 8e8, 3.0, 0.3
    pot = PotentialGPU(type="Disk", surfaceDensity=S, scaleRadius=Rd, scaleHeight=h)
    assert isinstance(pot, CompositePotentialGPU) and len(pot) == 2
    _, coefs = disk_potential_parts(surfaceDensity=S, scaleRadius=Rd, scaleHeight=h)
    orc = OracleMultipole.from_coefs(coefs)
    x = _pts(20000, 34, a=Rd)
    R = np.hypot(x[:, 0], x[:, 1])
    x = x[(R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))]
    Pa, ga, Ha, _ = A.disk_ansatz(x, S, Rd, h, 0.0)
    pm, Fm, Dm = orc.eval_deriv(x)
    _parity(pot.evalDeriv(torch.as_tensor(x, device=DEV)), (Pa + pm, Fm - ga, Dm - Ha))
    pts = np.array([[1.0, 0, 0.0], [3.0, 0, 0.0], [8.0, 0, 0.0], [3.0, 0, 0.3], [3.0, 0, 1.0],
                    [20.0, 0, 5.0]])
    p = pot.potential(torch.as_tensor(pts, device=DEV)).cpu().numpy()
    ref = np.array([_kuijken_gilmore(Rr, z, S, Rd, h) for Rr, _, z in pts])
    assert np.max(np.abs(p - ref) / np.abs(ref)) < 2e-5
