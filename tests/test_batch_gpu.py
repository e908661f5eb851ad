"""Batched ragged ensembles (agx_multipole_batch_eval) on gfx950.

An ensemble of potentials drawn like the 256-potential workload (lmax 4..12,
30..80 radial nodes, triaxial Dehnen tables) with particle counts scaled down
and edge counts added (0, 1, one more than an item) is evaluated in one launch
and compared with one plain launch per potential on the same rows (same
arithmetic: 1e-13 of the scale) and, for a few potentials, with the oracle
(tolerances of the Multipole parity tests).  Both paths (lane, LDS-staged)
and several item sizes are forced in turn; a hipGraph replay of the batched launch
must reproduce the eager result every time.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module")
def ensemble_parts():
    from agama_gpu_amd import MultipolePotentialGPU
    from agama_gpu_amd.workloads import ensemble_coefs, ensemble_spec, hernquist_points
    spec = ensemble_spec(24, seed=3)
    counts = [max(1, s["count"] // 40) for s in spec]
    counts[2], counts[5], counts[9] = 0, 1, 16385
    coefs = [ensemble_coefs(s) for s in spec]
    pots = [MultipolePotentialGPU(c, device=DEV) for c in coefs]
    x = hernquist_points(int(sum(counts)), 7)
    return coefs, pots, counts, x


def _rel(a, b):
    return (torch.max(torch.abs(a - b)) / torch.max(torch.abs(b))).item()


@pytest.mark.parametrize("path,rows", [(0, 4096), (0, 16384), (1, 2048), (1, 8192), (1, 16384)])
@pytest.mark.parametrize("order", ["pot", "pf", "deriv"])
def test_batched_equals_per_potential(ensemble_parts, path, rows, order):
    from agama_gpu_amd.batch import MultipoleEnsembleGPU
    _, pots, counts, x = ensemble_parts
    print(f"case path={path} rows={rows} order={order} n={len(x)}", flush=True)
    ens = MultipoleEnsembleGPU(pots, counts, choices=[(path, rows)] * len(pots))
    xyz = torch.as_tensor(x, device=DEV)
    n = len(x)
    mk = lambda *s: torch.full(s, np.nan, dtype=torch.float64, device=DEV)  # noqa: E731
    outs_b = (mk(n), mk(n, 3) if order != "pot" else None, mk(n, 6) if order == "deriv" else None)
    outs_s = (mk(n), mk(n, 3) if order != "pot" else None, mk(n, 6) if order == "deriv" else None)
    ens.eval_into(xyz, *outs_b)
    ens.eval_separately(xyz, *outs_s)
    for a, b in zip(outs_b, outs_s):
        if b is None:
            continue
        assert torch.isfinite(a).all(), "rows left unwritten"
        assert _rel(a, b) < 1e-13


@pytest.mark.parametrize("path,rows", [(0, 4096), (1, 8192)])
def test_batched_bad_positions(ensemble_parts, path, rows):
    """NaN coordinates give NaN outputs and r^2 = inf gives 0 in every
    potential's rows, as in one launch per potential; the rows around them
    agree with that launch as in the test above."""
    from agama_gpu_amd.batch import MultipoleEnsembleGPU
    _, pots, counts, x = ensemble_parts
    x = x.copy()
    rng = np.random.default_rng(17)
    bad = rng.choice(len(x), 64, replace=False)
    isnan = np.arange(64) % 2 == 0
    x[bad[isnan], rng.integers(0, 3, isnan.sum())] = np.nan
    x[bad[~isnan]] = [1e200, -1e200, 1e170] * np.sign(rng.standard_normal((32, 3)))
    print(f"case bad positions path={path} rows={rows}", flush=True)
    ens = MultipoleEnsembleGPU(pots, counts, choices=[(path, rows)] * len(pots))
    xyz = torch.as_tensor(x, device=DEV)
    b = [t.cpu().numpy() for t in ens.eval(xyz, der=True)]
    n = len(x)
    mk = lambda *s: torch.full(s, 7.0, dtype=torch.float64, device=DEV)  # noqa: E731
    s = (mk(n), mk(n, 3), mk(n, 6))
    ens.eval_separately(xyz, *s)
    s = [t.cpu().numpy() for t in s]
    good = np.ones(n, bool)
    good[bad] = False
    for a, r in zip(b, s):
        a2, r2 = a.reshape(n, -1), r.reshape(n, -1)
        assert np.all(np.isnan(a2[bad[isnan]]).any(1)) and np.all(np.isnan(r2[bad[isnan]]).any(1))
        assert np.all(a2[bad[~isnan]] == 0.0) and np.all(r2[bad[~isnan]] == 0.0)
        assert np.max(np.abs(a2[good] - r2[good])) <= 1e-13 * np.max(np.abs(r2[good]))


def test_batched_against_oracle(ensemble_parts):
    from agama_gpu_amd.batch import MultipoleEnsembleGPU
    from oracle.oracle import OracleMultipole
    coefs, pots, counts, x = ensemble_parts
    ens = MultipoleEnsembleGPU(pots, counts)
    xyz = torch.as_tensor(x, device=DEV)
    phi, F, _ = ens.eval(xyz)
    phi, F = phi.cpu().numpy(), F.cpu().numpy()
    for p in (0, 4, 9, 13, 23):
        sl = ens.slice(p)
        if sl.stop == sl.start:
            continue
        orc = OracleMultipole.from_coefs(coefs[p])
        po, Fo = orc.pot_force(x[sl], 4)
        assert np.max(np.abs(phi[sl] - po) / np.abs(po)) < 1e-12
        assert np.max(np.abs(F[sl] - Fo)) / np.max(np.abs(Fo)) < 1e-11


def test_batched_graph_replay(ensemble_parts):
    from agama_gpu_amd.batch import MultipoleEnsembleGPU
    _, pots, counts, x = ensemble_parts
    ens = MultipoleEnsembleGPU(pots, counts)
    xyz = torch.as_tensor(x, device=DEV)
    n = len(x)
    phi = torch.empty(n, dtype=torch.float64, device=DEV)
    F = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    ens.eval_into(xyz, phi, F)
    torch.cuda.synchronize()
    ref_phi, ref_F = phi.clone(), F.clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ens.eval_into(xyz, phi, F)          # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ens.eval_into(xyz, phi, F)
    for rep in range(3):
        phi.fill_(0.0)
        F.fill_(0.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(phi, ref_phi), f"replay {rep}"
        assert torch.equal(F, ref_F), f"replay {rep}"


def test_autotune_choice_is_feasible_and_correct(ensemble_parts):
    from agama_gpu_amd import autotune
    from agama_gpu_amd.batch import MultipoleEnsembleGPU
    _, pots, counts, x = ensemble_parts
    sub = [0, 1, 3]
    ch = autotune.tune_ensemble([pots[j] for j in sub])
    for j, (path, rows) in zip(sub, ch):
        assert (path, rows) in autotune.candidates(pots[j])
    ens = MultipoleEnsembleGPU(pots, counts, tune=True)
    xyz = torch.as_tensor(x, device=DEV)
    a = ens.eval(xyz)[1]
    b = torch.empty_like(a)
    ens.eval_separately(xyz, None, b)
    assert _rel(a, b) < 1e-13
