"""The differentiable table rebuild (torch) reproduces the kernel tables of
multipole_prep.build_multipole_data (numpy / scipy banded solve), and the
spline centre's knot weights reproduce ShiftedPotentialGPU's trajectory.
CPU only: no kernel is launched."""
import numpy as np
import pytest
import torch

from conftest import FIXTURES


def _coefs(name):
    from agama_gpu_amd.coefs import read_coefs
    from agama_gpu_amd.workloads import halo_lmax8_coefs
    if name == "halo_lmax8":
        return halo_lmax8_coefs()
    return read_coefs(FIXTURES / {"fire_dense": "600.dark.none_8.coef_mul_DR",
                                  "lmc": "100.LMC.none_8.coef_mult"}[name])


@pytest.mark.parametrize("prune", [0.0, 1e-16])
@pytest.mark.parametrize("name", ["fire_dense", "lmc", "halo_lmax8"])
def test_tables_match_numpy_build(name, prune):
    from agama_gpu_amd.autograd import FittableMultipole
    from agama_gpu_amd.multipole_prep import build_multipole_data
    c = _coefs(name)
    d = build_multipole_data(c, prune_threshold=prune)
    m = FittableMultipole(c, prune_threshold=prune, device="cpu")
    with torch.no_grad():
        poly, scal = m.tables()
    poly = poly.numpy()
    ref = d["poly_km"]
    assert poly.shape == ref.shape
    # per-term scale: the largest coefficient of that column
    scale = np.maximum(np.abs(ref).max(axis=(0, 2)), 1e-300)[None, :, None]
    assert np.max(np.abs(poly - ref) / scale) < 1e-11
    want = np.array([d["invPhi0"], d["inner_U"], d["inner_W"], d["outer_U"], d["outer_W"]])
    assert np.allclose(scal.numpy(), want, rtol=1e-13, atol=0.0)


def test_tables_gradient_is_finite_and_local():
    """Autograd through the rebuild: d(poly)/d(phi entry) is nonzero only on
    that column (the quintic system couples radii, not harmonics), except the
    monopole, which enters every column through log-scaling."""
    from agama_gpu_amd.autograd import FittableMultipole
    c = _coefs("halo_lmax8")
    m = FittableMultipole(c, device="cpu")
    poly, _ = m.tables()
    col = 3                                   # a non-monopole kernel column
    g = torch.autograd.grad(poly[:, col, :].sum(), m.phi_table)[0].numpy()
    src = int(m.cols[col])
    assert np.isfinite(g).all()
    nz = np.nonzero(np.abs(g).max(axis=0))[0]
    assert set(nz.tolist()) <= {int(m.cols[0]), src}
    assert np.abs(g[:, src]).max() > 0


def test_spline_center_weights():
    from agama_gpu_amd.autograd import SplineCenter
    from agama_gpu_amd.composite import _linear_extrapolating_spline
    rng = np.random.default_rng(3)
    times = np.sort(rng.uniform(-2.0, 2.0, 9))
    knots = rng.normal(size=(9, 3)) * 50.0
    sc = SplineCenter(times, knots, device="cpu")
    f = _linear_extrapolating_spline(times, knots)
    for t in [-3.0, times[0], -0.3, 0.0, 1.1, times[-1], 2.7]:
        got = sc(t).detach().numpy()
        assert np.allclose(got, f(t), rtol=1e-13, atol=1e-12)
        # the weights are exactly the derivative of c(t) w.r.t. the knots
        w = sc.weights(t).numpy()
        g = torch.autograd.grad(sc(t)[1], sc.knots)[0].numpy()
        assert np.allclose(g[:, 1], w, rtol=0, atol=1e-15)
        assert np.allclose(g[:, [0, 2]], 0.0)


@pytest.mark.parametrize("name", ["bar", "disk"])
def test_cylspline_tables_match_numpy_build(name):
    """The CylSpline rebuild (mirroring, log-scaling, spline slopes by matrix,
    outer fit by pseudo-inverse) reproduces cylspline_prep's node data and W:
    the reference's full-z bar fixture and the half-space disk table."""
    from agama_gpu_amd.autograd import FittableCylSpline
    from agama_gpu_amd.coefs import read_coefs
    from agama_gpu_amd.cylspline_prep import build_cylspline_data
    from agama_gpu_amd.workloads import mw_disk_bar_coefs
    c = read_coefs(FIXTURES / "600.bar.none_8.coef_cylsp_DR") if name == "bar" else \
        mw_disk_bar_coefs()
    d = build_cylspline_data(c)
    m = FittableCylSpline(c, device="cpu")
    with torch.no_grad():
        nodes, W = m.tables()
    nodes, ref = nodes.numpy(), d["node_arr"]
    assert nodes.shape == ref.shape
    # per (harmonic, quantity) scale
    scale = np.maximum(np.abs(ref).max(axis=(1, 2)), 1e-300)[:, None, None, :]
    assert np.max(np.abs(nodes - ref) / scale) < 1e-11
    Wref = d["W_outer"]
    assert np.max(np.abs(W.numpy() - Wref)) <= 1e-10 * np.abs(Wref).max()
    # round trip through the coefficient set
    back = m.to_coefs()
    for k, mv in enumerate(m.m_values):
        assert np.array_equal(back.phi[mv], np.asarray(c.phi[mv]))


def test_cylspline_table_gradient_is_local_in_m():
    """d(nodes of harmonic k)/d(table) touches harmonic k and, through the
    log-scaling ratio, the m=0 table only."""
    from agama_gpu_amd.autograd import FittableCylSpline
    from agama_gpu_amd.coefs import read_coefs
    c = read_coefs(FIXTURES / "600.bar.none_8.coef_cylsp_DR")
    m = FittableCylSpline(c, device="cpu")
    nodes, _ = m.tables()
    k = 3
    g = torch.autograd.grad(nodes[k].sum(), m.phi_table)[0].numpy()
    assert np.isfinite(g).all()
    touched = set(np.nonzero(np.abs(g).max(axis=(1, 2)))[0].tolist())
    assert touched <= {0, k} and k in touched


@pytest.mark.parametrize("n", [100, 4096, 10007])
def test_shift_backward_centre_sum(n):
    """The shift's centre gradient (blocked column sum) is -sum_i g_i."""
    from agama_gpu_amd.autograd import _Shift
    rng = np.random.default_rng(n)
    x = torch.as_tensor(rng.normal(size=(n, 3)), dtype=torch.float64)
    c = torch.as_tensor(rng.normal(size=3), dtype=torch.float64).requires_grad_(True)
    g = torch.as_tensor(rng.normal(size=(n, 3)), dtype=torch.float64)
    out = _Shift.apply(x, c)
    assert torch.equal(out, x - c.detach()[None, :])
    (gc,) = torch.autograd.grad(out, c, g)
    assert torch.allclose(gc, -g.sum(0), rtol=1e-12, atol=1e-12)
