"""GPU parity of the fused composite kernel: every analytic type, mixtures,
Shifted / Scaled / Evolving wrappers and Multipole leaves inside one launch,
against the oracles (analytic: oracle/analytic_oracle.py, Multipole:
oracle/mp_oracle.c).  Tolerances (FP64 both sides, same formulas):
phi max|d|/|phi| <= 1e-11 (the analytic bar; pow() near r/(r+a) -> 1
cancels ~log10(r/a) digits); force, deriv, rho max|d|/max|ref| <= 1e-11."""
import numpy as np
import pytest
import torch

from conftest import FIXTURES
from oracle import analytic_oracle as A

pytestmark = pytest.mark.gpu
TOL_P, TOL_G = 1e-11, 1e-11

CASES = [
    ("NFW", dict(mass=1e12, scaleRadius=20.0)),
    ("Plummer", dict(mass=1e10, scaleRadius=0.5)),
    ("Hernquist", dict(mass=5e11, scaleRadius=10.0)),
    ("Dehnen", dict(mass=1e11, scaleRadius=3.0, gamma=0.5)),
    ("Dehnen", dict(mass=1e11, scaleRadius=3.0, gamma=2.0)),
    ("Isochrone", dict(mass=1e11, scaleRadius=2.0)),
    ("MiyamotoNagai", dict(mass=6e10, scaleRadius=3.0, scaleHeight=0.3)),
    ("LogHalo", dict(velocity=200.0, coreRadius=1.0, axisRatioY=0.9, axisRatioZ=0.7)),
    ("DiskAnsatz", dict(surfaceDensity=8e8, scaleRadius=3.0, scaleHeight=0.4, innerCutoffRadius=0.5)),
    ("DiskAnsatz", dict(surfaceDensity=8e8, scaleRadius=3.0, scaleHeight=-0.4)),
    ("DiskAnsatz", dict(surfaceDensity=8e8, scaleRadius=3.0, scaleHeight=0.0)),
    ("UniformAcceleration", dict(ax=1.0, ay=-2.0, az=0.5)),
]


def pts(n=20000, seed=1):
    from agama_gpu_amd.workloads import hernquist_points
    x = hernquist_points(n, seed)
    x[:3] = [[0.0, 0.0, 3.0], [2.0, -1.0, 0.0], [1e-3, 2e-3, -4.0]]
    return x


def cmp(gpu_out, ref, what):
    p, f, d = [t.cpu().numpy() for t in gpu_out]
    P, g, h, _ = ref
    assert np.max(np.abs(p - P) / (np.abs(P) + 1e-300)) < TOL_P, what
    assert np.max(np.abs(f + g)) <= TOL_G * max(np.max(np.abs(g)), 1e-300), what
    assert np.max(np.abs(d + h)) <= TOL_G * max(np.max(np.abs(h)), 1e-300), what


@pytest.mark.parametrize("name,kw", CASES, ids=[f"{c[0]}{i}" for i, c in enumerate(CASES)])
def test_analytic_types(name, kw):
    from agama_gpu_amd import PotentialGPU
    print(f"analytic {name} {kw}", flush=True)
    x = pts()
    pot = PotentialGPU(type=name, **kw)
    ref = A.evaluate(name, x, **kw)
    xt = torch.as_tensor(x, device="cuda")
    cmp(pot.evalDeriv(xt), ref, name)
    rho = pot.density(xt).cpu().numpy()
    assert np.max(np.abs(rho - ref[3])) <= TOL_G * max(np.max(np.abs(ref[3])), 1e-300)
    f = pot.force(xt).cpu().numpy()
    assert np.max(np.abs(f + ref[1])) <= TOL_G * max(np.max(np.abs(ref[1])), 1e-300)


def test_composite_shift_scale_evolving():
    from agama_gpu_amd import (CompositePotentialGPU, EvolvingPotentialGPU, MultipolePotentialGPU,
                               PotentialGPU, ScaledPotentialGPU, ShiftedPotentialGPU, read_coefs)
    from oracle.oracle import OracleMultipole
    x = pts(30000, 2)
    xt = torch.as_tensor(x, device="cuda")
    coefs = read_coefs(FIXTURES / "100.LMC.none_8.coef_mult")
    mp = MultipolePotentialGPU(coefs)
    orc = OracleMultipole.from_coefs(coefs)

    def mp_ref(y):
        p, F, D = orc.eval_deriv(y)
        return p, -F, -D, orc.density(y)

    nfw = PotentialGPU(type="NFW", mass=1e12, scaleRadius=20.0)
    plum = PotentialGPU(type="Plummer", mass=1e10, scaleRadius=0.5)
    # Shifted(Scaled(Multipole)) + NFW + Plummer, evaluated in one launch
    lmc = ShiftedPotentialGPU(ScaledPotentialGPU(mp, scale=2.0, ampl=0.5), center=[30.0, -10.0, 5.0])
    comp = nfw + plum + lmc
    assert isinstance(comp, CompositePotentialGPU) and len(comp) == 3
    refs = [A.nfw(x, 1e12, 20.0), A.plummer(x, 1e10, 0.5),
            A.transformed(mp_ref, x, ampl=0.5 * 0.5, sigma=0.5, tau=(30.0, -10.0, 5.0))]
    tot = [sum(r[k] for r in refs) for k in range(4)]
    cmp(comp.evalDeriv(xt), tot, "composite")
    rho = comp.density(xt).cpu().numpy()
    assert np.max(np.abs(rho - tot[3])) <= 1e-10 * np.max(np.abs(tot[3]))

    # Evolving: lerp between two snapshots, nearest mode, clamping
    snaps = [PotentialGPU(type="Hernquist", mass=5e11, scaleRadius=10.0),
             PotentialGPU(type="Hernquist", mass=7e11, scaleRadius=12.0)]
    ev = EvolvingPotentialGPU(snaps, [0.0, 2.0], interpolate=True)
    r0, r1 = A.hernquist(x, 5e11, 10.0), A.hernquist(x, 7e11, 12.0)
    for t, w in ((0.5, 0.25), (-1.0, 0.0), (3.0, 1.0)):
        ref = [(1 - w) * a + w * b for a, b in zip(r0, r1)]
        cmp(ev.evalDeriv(xt, t=t), ref, f"evolving t={t}")
    evn = EvolvingPotentialGPU(snaps, [0.0, 2.0], interpolate=False)
    cmp(evn.evalDeriv(xt, t=1.2), r1, "evolving nearest")

    # time-dependent centre: (T,4) trajectory, cubic spline + linear extrapolation
    T = np.linspace(0.0, 10.0, 11)
    traj = np.stack([T, 5.0 * np.sin(T / 3), 2.0 * T, -T ** 2 / 10], axis=1)
    sh = ShiftedPotentialGPU(nfw, center=traj)
    from scipy.interpolate import CubicSpline
    cs = CubicSpline(T, traj[:, 1:], bc_type="not-a-knot")
    for t in (3.3, 12.0):
        c = cs(t) if t <= 10 else cs(10.0) + cs(10.0, 1) * (t - 10.0)
        cmp(sh.evalDeriv(xt, t=t), A.transformed(lambda y: A.nfw(y, 1e12, 20.0), x, tau=c),
            f"shifted t={t}")


def test_many_leaves_chunked():
    """More leaves than one launch takes: chunks accumulate."""
    from agama_gpu_amd import CompositePotentialGPU, PotentialGPU
    x = pts(5000, 4)
    comps = [PotentialGPU(type="Plummer", mass=1e9 * (i + 1), scaleRadius=0.5 + 0.1 * i)
             for i in range(21)]
    comp = CompositePotentialGPU(comps)
    p, f, d = comp.evalDeriv(torch.as_tensor(x, device="cuda"))
    ref = [sum(A.plummer(x, 1e9 * (i + 1), 0.5 + 0.1 * i)[k] for i in range(21)) for k in range(4)]
    cmp((p, f, d), ref, "chunked")


def test_factory_ini(tmp_path):
    from agama_gpu_amd import PotentialGPU
    ini = tmp_path / "mw.ini"
    (tmp_path / "lmc.coef").write_text((FIXTURES / "100.LMC.none_8.coef_mult").read_text())
    ini.write_text("[Potential halo]\ntype=NFW\nmass=1e12\nscaleRadius=20\n\n"
                   "[Potential disk]\ntype=MiyamotoNagai\nmass=6e10\nscaleRadius=3\nscaleHeight=0.3\n\n"
                   "[Potential lmc]\ntype=Multipole\nfile=lmc.coef\n")
    pot = PotentialGPU(file=str(ini))
    x = pts(2000, 5)
    p = pot.potential(torch.as_tensor(x, device="cuda")).cpu().numpy()
    from agama_gpu_amd import MultipolePotentialGPU, read_coefs
    lmc = MultipolePotentialGPU(read_coefs(FIXTURES / "100.LMC.none_8.coef_mult"))
    ref = A.nfw(x, 1e12, 20.0)[0] + A.miyamoto_nagai(x, 6e10, 3.0, 0.3)[0] + \
        lmc.potential(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert np.max(np.abs(p - ref) / np.abs(ref)) < 1e-12


@pytest.mark.parametrize("n", [1, 255, 257, 2_500_001])
def test_analytic_tiles(n):
    """Analytic-only sums run a kernel that moves 256-particle tiles through
    LDS with a grid of at most 8192 workgroups: sizes below one tile, one
    past a tile, and beyond 8192 tiles (some workgroups take two, the last
    one partial) for every output (phi, force, deriv, density), with two
    leaves (NFW + shifted Plummer), against the oracle."""
    from agama_gpu_amd import PotentialGPU
    print(f"analytic tiles n={n}", flush=True)
    from agama_gpu_amd.workloads import hernquist_points
    x = hernquist_points(max(n, 8), 7)[:n]
    nfw = dict(mass=1e12, scaleRadius=20.0)
    plu = dict(mass=1e10, scaleRadius=0.5)
    pot = PotentialGPU(type="NFW", **nfw) + PotentialGPU(type="Plummer", center=[1.0, -2.0, 0.5],
                                                        **plu)
    c = np.array([1.0, -2.0, 0.5])
    r1, r2 = A.evaluate("NFW", x, **nfw), A.evaluate("Plummer", x - c, **plu)
    ref = [r1[k] + r2[k] for k in range(4)]
    xt = torch.as_tensor(x, device="cuda")
    cmp(pot.evalDeriv(xt), ref, f"evalDeriv n={n}")
    phi, F = pot.eval(xt, pot=True, acc=True)
    assert np.max(np.abs(phi.cpu().numpy() - ref[0]) / np.abs(ref[0])) < TOL_P
    assert np.max(np.abs(F.cpu().numpy() + ref[1])) <= TOL_G * np.max(np.abs(ref[1]))
    p = pot.potential(xt).cpu().numpy()
    assert np.max(np.abs(p - ref[0]) / np.abs(ref[0])) < TOL_P
    rho = pot.density(xt).cpu().numpy()
    assert np.max(np.abs(rho - ref[3])) <= TOL_G * np.max(np.abs(ref[3]))
