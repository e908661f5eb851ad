"""The parity configurations at reduced N, against the oracle.

* Composite Milky-Way-like model, evalDeriv in one fused launch: Multipole
  halo (lmax 8 triaxial Dehnen), CylSpline disk + bar (Miyamoto-Nagai + m=2,4,6
  bar, 32 x 32 grid), NFW (1e12, 20 kpc), Plummer bulge (1e10, 0.5 kpc);
  20k Hernquist points.  Tolerances of the component parity tests: phi 1e-12
  relative, force 1e-11 and deriv 1e-10 of the maximum (deriv inside the
  CylSpline grid, away from the z-axis conditioning band).
* High-order Multipole: lmax 32, 80 nodes, all harmonics (1085 active),
  evalDeriv at 20k points log-uniform in r over [0.05, 800] kpc; every
  kernel path.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def test_mw_composite_evalderiv():
    from agama_gpu_amd.workloads import halo_lmax8_coefs, hernquist_points, mw_composite, \
        mw_disk_bar_coefs
    from oracle import analytic_oracle as A
    from oracle.cylspline_oracle import OracleCylSpline
    from oracle.oracle import OracleMultipole
    print("case mw composite evalDeriv n=20000", flush=True)
    comp = mw_composite(device=DEV)
    x = hernquist_points(20000, 11)
    p, F, D = [t.cpu().numpy() for t in comp.evalDeriv(torch.as_tensor(x, device=DEV))]
    pm, Fm, Dm = OracleMultipole.from_coefs(halo_lmax8_coefs()).eval_deriv(x)
    Pc, gc, Hc, outer = OracleCylSpline(mw_disk_bar_coefs()).eval(x)
    Pn, gn, Hn, _ = A.nfw(x, 1e12, 20.0)
    Pp, gp, Hp, _ = A.plummer(x, 1e10, 0.5)
    P = pm + Pc + Pn + Pp
    Fr = Fm - gc - gn - gp
    Dr = Dm - Hc - Hn - Hp
    assert np.max(np.abs(p - P) / np.abs(P)) < 1e-12
    assert np.max(np.abs(F - Fr)) <= 1e-11 * np.abs(Fr).max()
    R = np.hypot(x[:, 0], x[:, 1])
    ok = ((R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))) & ~outer
    assert ok.mean() > 0.5
    assert np.max(np.abs(D[ok] - Dr[ok])) <= 1e-10 * np.abs(Dr[ok]).max()


@pytest.fixture(scope="module")
def high_order():
    from agama_gpu_amd.workloads import high_order_coefs, loguniform_points
    from oracle.oracle import OracleMultipole
    c = high_order_coefs()
    x = loguniform_points(20000, 2, 0.05, 800.0)
    return c, x, OracleMultipole.from_coefs(c).eval_deriv(x, 8)


@pytest.mark.parametrize("path", [0, 1, 2, 3, 4])
def test_high_order_lmax32(high_order, path):
    from agama_gpu_amd import MultipolePotentialGPU, _lib
    c, x, (po, Fo, Do) = high_order
    print(f"case lmax32 path={path} n={len(x)}", flush=True)
    pot = MultipolePotentialGPU(c, device=DEV)
    assert pot.lmax == 32 and pot.n_lm > 1000
    _lib.set_option(_lib.AGX_OPT_MP_PATH, path)
    try:
        p, F, D = [t.cpu().numpy() for t in pot.evalDeriv(torch.as_tensor(x, device=DEV))]
        pf = pot.force(torch.as_tensor(x, device=DEV)).cpu().numpy()
    finally:
        _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)
    assert np.max(np.abs(p - po) / np.abs(po)) < 1e-12
    assert np.max(np.abs(F - Fo)) <= 1e-11 * np.abs(Fo).max()
    assert np.max(np.abs(pf - Fo)) <= 1e-11 * np.abs(Fo).max()
    R = np.hypot(x[:, 0], x[:, 1])
    ok = (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))
    assert np.max(np.abs(D[ok] - Do[ok])) <= 1e-10 * np.abs(Do[ok]).max()


def test_high_order_lmax32_sorted_auto():
    """The benchmarked form of the lmax-32 workload: at N >= 200k the automatic
    choice for a table too large for LDS (4.1 MB) is the interval-sorted path
    (agx_multipole_eval_sorted); 250k points log-uniform in r over
    [0.05, 800] kpc, evalDeriv and pot+force, against the oracle."""
    from agama_gpu_amd import MultipolePotentialGPU, _lib
    from agama_gpu_amd.multipole import eval_table
    from agama_gpu_amd.workloads import high_order_coefs, loguniform_points
    from oracle.oracle import OracleMultipole
    c = high_order_coefs()
    x = loguniform_points(250_000, 12, 0.05, 800.0)
    print(f"case lmax32 sorted auto n={len(x)}", flush=True)
    pot = MultipolePotentialGPU(c, device=DEV)
    assert _lib.get_option(_lib.AGX_OPT_MP_PATH) == 0
    xt = torch.as_tensor(x, device=DEV)
    n = len(x)
    phi = torch.empty(n, dtype=torch.float64, device=DEV)
    F = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    D = torch.empty((n, 6), dtype=torch.float64, device=DEV)
    work = eval_table(pot._desc, pot._data, xt, phi, F, D, _lib.AGX_PHI | _lib.AGX_FORCE |
                      _lib.AGX_DERIV, xt.device)
    assert work is not None                     # the sorted path ran
    p, Fg, Dg = phi.cpu().numpy(), F.cpu().numpy(), D.cpu().numpy()
    po, Fo, Do = OracleMultipole.from_coefs(c).eval_deriv(x)
    assert np.max(np.abs(p - po) / np.abs(po)) < 1e-12
    assert np.max(np.abs(Fg - Fo)) <= 1e-11 * np.abs(Fo).max()
    R = np.hypot(x[:, 0], x[:, 1])
    ok = (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))
    assert np.max(np.abs(Dg[ok] - Do[ok])) <= 1e-10 * np.abs(Do[ok]).max()
    p2, F2 = [t.cpu().numpy() for t in pot.eval(xt, pot=True, acc=True)]
    assert np.max(np.abs(p2 - po) / np.abs(po)) < 1e-12
    assert np.max(np.abs(F2 - Fo)) <= 1e-11 * np.abs(Fo).max()
