"""GPU parity on random Multipole table structures (tests/random_tables.py).

Eighteen seeded tables (lmax 0..12; sin terms, odd l and m, sparse subsets,
even-even programs, sign-changing monopoles without log-scaling) at points
spread over both extrapolation regions, near and on the z axis, at the origin
and at r_max, through every evaluation path of the library: per-lane, binned,
LDS-staged, interval-sorted, the fused composite kernel (as a shifted and
scaled leaf) and the ragged-ensemble launch.  Compared with the oracle at the
tolerances of the fixture tests (global scale: a sign-changing monopole has
zeros):
  phi 1e-12, force 1e-11, deriv 1e-10 (rows with 0 < sin(theta) < 1e-6
  excluded, as there), density 1e-10.
"""
import numpy as np
import pytest
import torch

from random_tables import random_multipole, sweep_points

pytestmark = pytest.mark.gpu

SEEDS = list(range(18))
N = 20_000


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def regular(x):
    R = np.hypot(x[:, 0], x[:, 1])
    return (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))


@pytest.fixture(scope="module")
def tables():
    from oracle.oracle import OracleMultipole
    out = {}
    for s in SEEDS:
        c, d = random_multipole(s)
        x = sweep_points(d, N, s)
        o = OracleMultipole.from_coefs(c)
        out[s] = (c, d, x, o.eval_deriv(x), o.density(x))
    return out


@pytest.fixture(params=[1, 2, 3, 4], ids=["lane", "binned", "lds", "sorted"])
def mp_path(request):
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_MP_PATH, request.param)
    yield request.param
    _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)


@pytest.mark.parametrize("seed", SEEDS)
def test_random_table_paths(tables, seed, mp_path):
    from agama_gpu_amd import MultipolePotentialGPU
    c, d, x, (po, Fo, Do), rho_o = tables[seed]
    print(f"case seed={seed} lmax={d['lmax']} pattern={d['pattern']} signed={d['signed']} "
          f"nR={d['nR']} path={mp_path} n={len(x)}", flush=True)
    gpu = MultipolePotentialGPU(c)
    xt = torch.as_tensor(x, device="cuda")
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(xt)]
    assert rel(p, po) < 1e-12, d
    assert rel(F, Fo) < 1e-11, d
    ok = regular(x)
    assert rel(D[ok], Do[ok]) < 1e-10, d
    assert np.all(np.isfinite(D)), d
    p2, F2 = [t.cpu().numpy() for t in gpu.eval(xt, pot=True, acc=True)]
    assert rel(p2, po) < 1e-12 and rel(F2, Fo) < 1e-11, d
    assert rel(gpu.potential(xt).cpu().numpy(), po) < 1e-12, d
    assert rel(gpu.density(xt).cpu().numpy(), rho_o) < 1e-10, d


@pytest.mark.parametrize("seed", SEEDS[::2])
def test_random_table_as_composite_leaf(tables, seed):
    """Shifted by c and scaled by 2 with amplitude 0.7 inside a composite:
    Phi = 0.7 * s * Phi_mp(s (x - c)), s = 1/2."""
    from agama_gpu_amd import (CompositePotentialGPU, MultipolePotentialGPU, ScaledPotentialGPU,
                               ShiftedPotentialGPU)
    from oracle.oracle import OracleMultipole
    c, d, x, _, _ = tables[seed]
    print(f"case seed={seed} composite leaf", flush=True)
    ctr = np.array([0.3, -0.2, 0.1]) * d["r_min"]
    pot = CompositePotentialGPU([ShiftedPotentialGPU(
        ScaledPotentialGPU(MultipolePotentialGPU(c), 2.0, ampl=0.7), ctr)])
    xt = torch.as_tensor(x, device="cuda")
    p, F, D = [t.cpu().numpy() for t in pot.evalDeriv(xt)]
    s = 0.5
    po, Fo, Do = OracleMultipole.from_coefs(c).eval_deriv(s * (x - ctr))
    po, Fo, Do = 0.7 * s * po, 0.7 * s * s * Fo, 0.7 * s ** 3 * Do
    assert rel(p, po) < 1e-12, d
    assert rel(F, Fo) < 1e-11, d
    ok = regular(x - ctr)
    assert rel(D[ok], Do[ok]) < 1e-10, d


def test_random_tables_as_one_ensemble(tables):
    """All eighteen tables in one ragged-ensemble launch (each with its own
    rows) against the oracle."""
    from agama_gpu_amd import MultipolePotentialGPU
    from agama_gpu_amd.batch import MultipoleEnsembleGPU
    pots, xs, counts = [], [], []
    for s in SEEDS:
        c, d, x, _, _ = tables[s]
        k = 2000 + 611 * s
        pots.append(MultipolePotentialGPU(c))
        xs.append(x[:k])
        counts.append(k)
    x = np.concatenate(xs)
    ens = MultipoleEnsembleGPU(pots, counts)
    p, F, _ = [None if t is None else t.cpu().numpy() for t in ens.eval(torch.as_tensor(x, device="cuda"))]
    for s in SEEDS:
        sl = ens.slice(s)
        po, Fo, _ = tables[s][3]
        k = sl.stop - sl.start
        assert rel(p[sl], po[:k]) < 1e-12, tables[s][1]
        assert rel(F[sl], Fo[:k]) < 1e-11, tables[s][1]
