"""Pins the analytic oracle (oracle/analytic_oracle.py) by calculus and closed
forms: grad = FD of Phi, Hessian = FD of grad (slot order
[xx, yy, zz, xy, yz, xz]), density = trace(Hessian)/(4 pi G) (Poisson), and
known special values (central potentials, limits that reduce one model to
another)."""
import math

import numpy as np
import pytest

from oracle import analytic_oracle as A

G = A.G
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
    ("DiskAnsatz", dict(surfaceDensity=8e8, scaleRadius=3.0, scaleHeight=-0.4, innerCutoffRadius=0.0)),
    ("UniformAcceleration", dict(ax=1.0, ay=-2.0, az=0.5)),
]


def pts(n=300, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    return u * np.exp(rng.uniform(np.log(0.2), np.log(80.0), n))[:, None]


@pytest.mark.parametrize("name,kw", CASES, ids=[f"{c[0]}{i}" for i, c in enumerate(CASES)])
def test_calculus(name, kw):
    x = pts()
    # keep away from z = 0 where the disk's |z| makes the Hessian jump
    x[:, 2] = np.where(np.abs(x[:, 2]) < 0.05, 0.05, x[:, 2])
    P, g, h, rho = A.evaluate(name, x, **kw)
    step = 1e-5 * np.linalg.norm(x, axis=1)[:, None]
    e = np.eye(3)
    fd = np.stack([(A.evaluate(name, x + step * e[k], **kw)[0] - A.evaluate(name, x - step * e[k], **kw)[0])
                   / (2 * step[:, 0]) for k in range(3)], axis=1)
    assert np.max(np.abs(fd - g)) <= 1e-6 * np.max(np.abs(g))
    gp = [A.evaluate(name, x + step * e[k], **kw)[1] for k in range(3)]
    gm = [A.evaluate(name, x - step * e[k], **kw)[1] for k in range(3)]
    scale = max(np.max(np.abs(h)), 1e-300)
    for s, (i, j) in enumerate(A.SLOTS):
        fdh = (gp[j][:, i] - gm[j][:, i]) / (2 * step[:, 0])
        assert np.max(np.abs(fdh - h[:, s])) <= 1e-5 * scale + 1e-300, (name, s)
    lap = (h[:, 0] + h[:, 1] + h[:, 2]) * A.INV_4PIG
    assert np.max(np.abs(lap - rho)) <= 1e-9 * max(np.max(np.abs(rho)), 1e-300)


@pytest.mark.filterwarnings("ignore::RuntimeWarning")
def test_special_values():
    o = np.zeros((1, 3))
    assert A.plummer(o, 1e10, 0.5)[0][0] == pytest.approx(-G * 1e10 / 0.5, rel=1e-15)
    assert A.isochrone(o, 1e11, 2.0)[0][0] == pytest.approx(-G * 1e11 / 4.0, rel=1e-15)
    assert A.nfw(o, 1e12, 20.0)[0][0] == pytest.approx(-G * 1e12 / 20.0, rel=1e-12)
    x = pts(50, 3)
    # Dehnen gamma=1 is Hernquist; Miyamoto-Nagai with a=0 is Plummer with b
    for a_, b_ in zip(A.dehnen(x, 3e11, 4.0, 1.0)[:3], A.hernquist(x, 3e11, 4.0)[:3]):
        assert np.max(np.abs(a_ - b_)) <= 1e-12 * np.max(np.abs(b_))
    for a_, b_ in zip(A.miyamoto_nagai(x, 3e11, 0.0, 1.5)[:3], A.plummer(x, 3e11, 1.5)[:3]):
        assert np.max(np.abs(a_ - b_)) <= 1e-12 * np.max(np.abs(b_))
    # flat rotation curve of a spherical log halo far outside the core
    r = np.array([[500.0, 0, 0]])
    g = A.loghalo(r, 200.0, 0.1, 1.0, 1.0)[1]
    assert math.sqrt(g[0, 0] * 500.0) == pytest.approx(200.0, rel=1e-6)
    # Hernquist: r^2 dPhi/dr -> G M
    g = A.hernquist(np.array([[1e7, 0, 0]]), 5e11, 10.0)[1]
    assert g[0, 0] * 1e14 == pytest.approx(G * 5e11, rel=1e-5)
    # the disk's vertical profile H'' integrates to 1 (surface density normalisation)
    for hz in (0.4, -0.4):
        zz = np.linspace(-40, 40, 400001)
        if hz > 0:
            H2 = 0.5 * np.exp(-np.abs(zz) / hz) / hz
        else:
            H2 = 0.5 / np.cosh(zz / -hz) ** 2 / -hz
        assert np.trapezoid(H2, zz) == pytest.approx(1.0, rel=1e-6)


def test_transform_chain_rule():
    """Shifted/Scaled transform: ampl f(sigma (x - tau)) derivatives vs FD."""
    f = lambda y: A.hernquist(y, 5e11, 10.0)
    x = pts(100, 8)
    P, g, h, _ = A.transformed(f, x, ampl=0.7, sigma=1.3, tau=(1.0, -2.0, 0.5))
    step = 1e-5 * np.linalg.norm(x, axis=1)[:, None]
    e = np.eye(3)
    fd = np.stack([(A.transformed(f, x + step * e[k], 0.7, 1.3, (1.0, -2.0, 0.5))[0] -
                    A.transformed(f, x - step * e[k], 0.7, 1.3, (1.0, -2.0, 0.5))[0]) / (2 * step[:, 0])
                   for k in range(3)], axis=1)
    assert np.max(np.abs(fd - g)) <= 1e-6 * np.max(np.abs(g))


@pytest.mark.parametrize("hz", [0.3, -0.3])
def test_disk_ansatz_midplane_against_mpmath(hz):
    """Near the midplane the vertical profiles grow as z^2; the oracle's
    Phi and dPhi/dz there equal a 50-digit mpmath evaluation of the closed
    forms h/2 (e^-u - 1 + u) and |h|/2 ln cosh u (u = |z|/|h|) to 1e-14
    (the reference's form e^-u - 1 + u keeps only 1e-16 / u^2)."""
    import mpmath as mp
    mp.mp.dps = 50
    Sigma, Rd = 5e8, 3.0
    z = np.concatenate([-np.geomspace(1e-9, 2.0, 40), np.geomspace(1e-9, 2.0, 40)])
    x = np.stack([np.full_like(z, 1.7), np.full_like(z, -0.4), z], axis=1)
    P, g, _, _ = A.disk_ansatz(x, Sigma, Rd, hz)
    for k in range(len(z)):
        X, Y, Z = (mp.mpf(float(v)) for v in x[k])
        r = mp.sqrt(X * X + Y * Y + Z * Z)
        f = 4 * mp.pi * mp.mpf(A.G) * Sigma * mp.exp(-r / Rd)
        h = abs(mp.mpf(hz))
        u = abs(Z) / h
        if hz > 0:
            Hv, H1 = h / 2 * (mp.exp(-u) - 1 + u), mp.sign(Z) * (1 - mp.exp(-u)) / 2
        else:
            Hv, H1 = h / 2 * mp.log(mp.cosh(u)), mp.tanh(Z / h) / 2
        fr = -f / Rd
        ref_P = f * Hv
        ref_gz = fr * Hv * Z / r + f * H1
        assert abs(P[k] - float(ref_P)) <= 1e-14 * abs(float(ref_P)), (hz, z[k])
        assert abs(g[k, 2] - float(ref_gz)) <= 1e-14 * abs(float(ref_gz)), (hz, z[k])
