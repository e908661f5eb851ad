"""GPU parity on random parameters of the analytic family and on random
wrapper trees, against the oracles.

* Every analytic type with six seeded parameter sets each (masses, scale
  radii and heights log-uniform over decades; Dehnen gamma including 0, 1 and
  2; DiskAnsatz with exponential, isothermal and thin vertical profiles, with
  and without an inner cutoff; LogHalo axis ratios in (0.5, 1]) at points
  from 1e-3 to 1e3 scale radii, on the z axis and in the midplane.
* Random trees of 2..24 leaves (analytic types, a random Multipole table, a
  random CylSpline table) under random Shifted / Scaled / Composite nesting
  and an Evolving pair, evaluated in the fused composite launch (chunked
  past 16 leaves).
Tolerances: phi 1e-11 of max(|phi|, 1e-6 max|phi|) (LogHalo and uniform
acceleration cross zero; for trees 1e-11 of the sum of the leaves' |phi|, the
scale of the rounding of a sum whose parts cancel), force, deriv and rho
1e-11 of their maximum
(deriv 1e-10 with expansion leaves, rows near the axis excluded as in the
expansion tests).
"""
import numpy as np
import pytest
import torch

from oracle import analytic_oracle as A

pytestmark = pytest.mark.gpu

TYPES = ["NFW", "Plummer", "Hernquist", "Dehnen", "Isochrone", "MiyamotoNagai", "LogHalo",
         "DiskAnsatz", "UniformAcceleration"]


def random_params(name, rng, k):
    lu = lambda a, b: 10.0 ** rng.uniform(a, b)   # noqa: E731
    if name in ("NFW", "Plummer", "Hernquist", "Isochrone"):
        return dict(mass=lu(8, 13), scaleRadius=lu(-1, 2))
    if name == "Dehnen":
        return dict(mass=lu(8, 13), scaleRadius=lu(-1, 2),
                    gamma=[0.0, 1.0, 2.0, 0.5, 1.5, rng.uniform(0, 1.99)][k % 6])
    if name == "MiyamotoNagai":
        return dict(mass=lu(9, 12), scaleRadius=lu(-0.5, 1.5), scaleHeight=lu(-2, 0.5))
    if name == "LogHalo":
        return dict(velocity=rng.uniform(80, 300), coreRadius=lu(-1, 1),
                    axisRatioY=rng.uniform(0.5, 1.0), axisRatioZ=rng.uniform(0.5, 1.0))
    if name == "DiskAnsatz":
        h = [lu(-1.5, 0), -lu(-1.5, 0), 0.0][k % 3]
        kw = dict(surfaceDensity=lu(7, 9.5), scaleRadius=lu(-0.3, 1), scaleHeight=h)
        if k % 2:
            kw["innerCutoffRadius"] = lu(-1, 0.3)
        return kw
    return dict(ax=rng.normal(), ay=rng.normal(), az=rng.normal())


def scale_of(kw):
    return kw.get("scaleRadius", kw.get("coreRadius", 1.0))


def points(scale, n, seed):
    rng = np.random.default_rng(seed)
    r = scale * np.exp(rng.uniform(np.log(1e-3), np.log(1e3), n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    x = u * r[:, None]
    x[:8, :2] = 0.0                 # on the z axis
    x[8:16, 2] = 0.0                # in the midplane
    return x


def check(gpu_out, ref, what, tol_d=1e-11, rows=None, pscale=None):
    """pscale: per-point scale of phi (a sum of parts' |phi| where parts may
    cancel: the rounding of the sum is relative to its terms)."""
    p, f, d = [t.cpu().numpy() for t in gpu_out]
    P, g, h = ref[0], ref[1], ref[2]
    if pscale is None:
        pscale = np.maximum(np.abs(P), 1e-6 * np.abs(P).max())
    pe = np.abs(p - P) / pscale
    assert np.max(pe) < 1e-11, what
    assert np.max(np.abs(f + g)) <= 1e-11 * max(np.abs(g).max(), 1e-300), what
    rows = slice(None) if rows is None else rows
    assert np.max(np.abs(d[rows] + h[rows])) <= tol_d * max(np.abs(h[rows]).max(), 1e-300), what


@pytest.mark.parametrize("k", range(6))
@pytest.mark.parametrize("name", TYPES)
def test_random_analytic_parameters(name, k):
    from agama_gpu_amd import PotentialGPU
    rng = np.random.default_rng(100 * TYPES.index(name) + k)
    kw = random_params(name, rng, k)
    print(f"case {name} {kw}", flush=True)
    x = points(scale_of(kw), 20000, k)
    pot = PotentialGPU(type=name, **kw)
    ref = A.evaluate(name, x, **kw)
    xt = torch.as_tensor(x, device="cuda")
    check(pot.evalDeriv(xt), ref, (name, kw))
    rho = pot.density(xt).cpu().numpy()
    assert np.max(np.abs(rho - ref[3])) <= 1e-11 * max(np.abs(ref[3]).max(), 1e-300), (name, kw)


def _tree(rng, depth, pool):
    """(gpu potential, reference function x -> (P, grad, Hessian, rho, S)) with
    S the sum of the leaves' |P| (the scale of the rounding of a sum whose
    parts cancel)."""
    from agama_gpu_amd import (CompositePotentialGPU, ScaledPotentialGPU, ShiftedPotentialGPU)
    r = rng.random()
    if depth > 0 and r < 0.25:
        parts = [_tree(rng, depth - 1, pool) for _ in range(int(rng.integers(2, 4)))]

        def comp(x, fs=[f for _, f in parts]):
            outs = [f(x) for f in fs]
            return tuple(sum(o[k] for o in outs) for k in range(5))
        return CompositePotentialGPU([p for p, _ in parts]), comp
    if depth > 0 and r < 0.5:
        inner, f = _tree(rng, depth - 1, pool)
        c = rng.normal(size=3) * 2.0
        return ShiftedPotentialGPU(inner, c), lambda x, f=f, c=c: f(x - c)
    if depth > 0 and r < 0.7:
        inner, f = _tree(rng, depth - 1, pool)
        sc, am = 10.0 ** rng.uniform(-0.5, 0.5), rng.uniform(0.3, 2.0)
        s = 1.0 / sc

        def scaled(x, f=f, s=s, am=am):
            o = f(s * x)
            return (am * s * o[0], am * s * s * o[1], am * s ** 3 * o[2], am * s ** 3 * o[3],
                    am * s * o[4])
        return ScaledPotentialGPU(inner, sc, ampl=am), scaled
    return pool[int(rng.integers(len(pool)))]


def _with_abs(ref):
    return tuple(ref) + (np.abs(ref[0]),)


@pytest.fixture(scope="module")
def leaf_pool():
    from agama_gpu_amd import CylSplinePotentialGPU, MultipolePotentialGPU, PotentialGPU
    from oracle.cylspline_oracle import OracleCylSpline
    from oracle.oracle import OracleMultipole
    from random_tables import random_cylspline, random_multipole
    rng = np.random.default_rng(77)
    pool = []
    for i, name in enumerate(TYPES):
        kw = random_params(name, rng, i)
        if name == "DiskAnsatz":
            kw["scaleHeight"] = abs(kw["scaleHeight"]) or 0.3    # smooth vertical profile
        pool.append((PotentialGPU(type=name, **kw),
                     lambda x, name=name, kw=kw: _with_abs(A.evaluate(name, x, **kw))))
    mc, _ = random_multipole(4, lmax=6, pattern="full")
    om = OracleMultipole.from_coefs(mc)

    def mp_ref(x):
        p, F, D = om.eval_deriv(x)
        return p, -F, -D, om.density(x), np.abs(p)
    pool.append((MultipolePotentialGPU(mc), mp_ref))
    cc, _ = random_cylspline(5)
    oc = OracleCylSpline(cc)

    def cyl_ref(x):
        P, g, H, outer = oc.eval(x)
        rho = np.where(outer, 0.0, (H[:, 0] + H[:, 1] + H[:, 2]) / (4 * np.pi * 4.30091727067736e-06))
        return P, g, H, rho, np.abs(P)
    pool.append((CylSplinePotentialGPU(cc), cyl_ref))
    return pool


@pytest.mark.parametrize("seed", range(8))
def test_random_wrapper_trees(leaf_pool, seed):
    from agama_gpu_amd import CompositePotentialGPU
    rng = np.random.default_rng(900 + seed)
    n_top = int(rng.integers(2, 9)) if seed < 6 else int(rng.integers(17, 25))
    parts = [_tree(rng, 2, leaf_pool) for _ in range(n_top)]
    pot = CompositePotentialGPU([p for p, _ in parts])
    print(f"case tree seed={seed} top={n_top} leaves={len(pot._leaves())}", flush=True)
    x = points(5.0, 20000, seed)
    outs = [f(x) for _, f in parts]
    ref = tuple(sum(o[k] for o in outs) for k in range(5))
    R = np.hypot(x[:, 0], x[:, 1])
    rows = R > 1e-3 * np.linalg.norm(x, axis=1)      # off the axis (shifted expansion leaves)
    pscale = ref[4]
    check(pot.evalDeriv(torch.as_tensor(x, device="cuda")), ref, seed, tol_d=1e-10, rows=rows,
          pscale=pscale)


def test_random_evolving_pair(leaf_pool):
    """Lerp between two random trees at three times, nearest snapshot, clamping."""
    from agama_gpu_amd import EvolvingPotentialGPU
    rng = np.random.default_rng(5)
    a, fa = _tree(rng, 2, leaf_pool)
    b, fb = _tree(rng, 2, leaf_pool)
    x = points(5.0, 20000, 3)
    ra, rb = fa(x), fb(x)
    xt = torch.as_tensor(x, device="cuda")
    rows = np.hypot(x[:, 0], x[:, 1]) > 1e-3 * np.linalg.norm(x, axis=1)
    ev = EvolvingPotentialGPU([a, b], [1.0, 3.0], interpolate=True)
    for t, w in ((1.5, 0.25), (0.0, 0.0), (9.0, 1.0)):
        ref = tuple((1 - w) * ra[k] + w * rb[k] for k in range(3))
        sc = (1 - w) * ra[4] + w * rb[4]
        check(ev.evalDeriv(xt, t=t), ref, f"t={t}", tol_d=1e-10, rows=rows, pscale=sc)
    evn = EvolvingPotentialGPU([a, b], [1.0, 3.0], interpolate=False)
    check(evn.evalDeriv(xt, t=2.2), rb, "nearest", tol_d=1e-10, rows=rows, pscale=rb[4])
