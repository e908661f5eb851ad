"""GPU parity of the gfx950 Multipole kernels against the oracle.

Compared on the same seeded inputs with the reference's error metrics
(tests/test_phase1_multipole.py:47-54 of the reference):
  phi:   max|d| / (|phi_oracle| + 1e-30)              <= 1e-12
  force: max|dF_component| / max|F_oracle|           <= 1e-11
  deriv: max|d| / max|deriv_oracle|                  <= 1e-10
  rho:   max|d| / max|rho_oracle|                    <= 1e-10
Both sides are FP64; the differences are summation order and libm ulps.
"""
import numpy as np
import pytest
import torch

from conftest import FIXTURES

pytestmark = pytest.mark.gpu

TOL_PHI, TOL_F, TOL_D, TOL_RHO = 1e-12, 1e-11, 1e-10, 1e-10


def rel_phi(a, b):
    return np.max(np.abs(a - b) / (np.abs(b) + 1e-30))


def rel_glob(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def cases():
    from agama_gpu_amd.coefs import read_coefs
    from agama_gpu_amd.workloads import halo_lmax8_coefs
    return {
        "fire_dense": read_coefs(FIXTURES / "600.dark.none_8.coef_mul_DR"),
        "lmc": read_coefs(FIXTURES / "100.LMC.none_8.coef_mult"),
        "halo_lmax8": halo_lmax8_coefs(),
    }


def points(n, seed=0):
    from agama_gpu_amd.workloads import hernquist_points
    x = hernquist_points(n, seed)
    # exact special cases: origin, z axis (both poles), grid boundaries
    x[:6] = [[0, 0, 0], [0, 0, 3.0], [0, 0, -3.0], [1e-12, 0, 7.0], [0, 0, 1e-3], [5.0, -2.0, 0]]
    return x


@pytest.fixture(params=[1, 2, 3, 4], ids=["lane", "binned", "lds", "sorted"])
def mp_path(request):
    """Force one evaluation path of the native library for the test."""
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_MP_PATH, request.param)
    yield request.param
    _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)


@pytest.mark.parametrize("name", ["fire_dense", "lmc", "halo_lmax8"])
def test_pot_force_deriv_density(cases, name, mp_path):
    from agama_gpu_amd import MultipolePotentialGPU
    from oracle.oracle import OracleMultipole
    c = cases[name]
    print(f"case={name} path={mp_path} n=100000", flush=True)
    x = points(100_000, 0)
    gpu = MultipolePotentialGPU(c)
    orc = OracleMultipole.from_coefs(c)
    xt = torch.as_tensor(x, device="cuda")
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(xt)]
    po, Fo, Do = orc.eval_deriv(x)
    assert rel_phi(p, po) < TOL_PHI
    assert rel_glob(F, Fo) < TOL_F
    # The spherical-to-Cartesian Hessian (reference :728-799) cancels terms of
    # size |dPhi/dtheta| / sin(theta)^3 near the z axis; for 0 < sin(theta) < 1e-6
    # it is ill-conditioned in any FP64 evaluation, so such rows are excluded.
    R = np.hypot(x[:, 0], x[:, 1])
    ok = (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))
    assert rel_glob(D[ok], Do[ok]) < TOL_D
    # pot-only and pot+force kernels agree with the full one
    p1 = gpu.potential(xt).cpu().numpy()
    p2, F2 = [t.cpu().numpy() for t in gpu.eval(xt, pot=True, acc=True)]
    assert rel_phi(p1, po) < TOL_PHI and rel_phi(p2, po) < TOL_PHI
    assert rel_glob(F2, Fo) < TOL_F
    rho = gpu.density(xt).cpu().numpy()
    rho_o = orc.density(x)
    assert rel_glob(rho, rho_o) < TOL_RHO


def test_single_point_and_numpy_input(cases):
    from agama_gpu_amd import MultipolePotentialGPU
    from oracle.oracle import OracleMultipole
    c = cases["fire_dense"]
    gpu = MultipolePotentialGPU(c)
    orc = OracleMultipole.from_coefs(c)
    x = np.array([3.0, -4.0, 1.5])
    p = gpu.potential(x)
    f = gpu.force(x)
    fo, do = gpu.forceDeriv(x)
    assert p.shape == () and f.shape == (3,) and do.shape == (6,)
    po, Fo = orc.pot_force(x[None])
    assert abs(p.item() - po[0]) < 1e-12 * abs(po[0])
    assert np.max(np.abs(f.cpu().numpy() - Fo[0])) < 1e-11 * np.abs(Fo).max()


def test_empty_and_ragged_sizes(cases, mp_path):
    from agama_gpu_amd import MultipolePotentialGPU
    from oracle.oracle import OracleMultipole
    gpu = MultipolePotentialGPU(cases["lmc"])
    orc = OracleMultipole.from_coefs(cases["lmc"])
    assert gpu.potential(np.zeros((0, 3))).shape == (0,)
    for n in (1, 63, 64, 65, 255, 257, 1000):
        x = points(max(n, 6), n)[:n]
        p, F = gpu.eval(torch.as_tensor(x, device="cuda"), pot=True, acc=True)
        po, Fo = orc.pot_force(x)
        assert rel_phi(p.cpu().numpy(), po) < TOL_PHI
        assert rel_glob(F.cpu().numpy(), Fo) < TOL_F


@pytest.mark.parametrize("name", ["fire_dense", "lmc", "halo_lmax8"])
def test_program_flag_variants_agree(cases, name):
    """The LDS-staged kernel variants for programs that recur by 1 only
    (AGX_PROG_STRIDE1), by 2 only (AGX_PROG_STRIDE2) or have no sin terms
    (AGX_PROG_COS_ONLY) give the same results as the generic one on the
    programs the flags describe: the generic step's c = cos * fma(cos, h, 1-h)
    and dc = -sin * fma(h, cos, u) are exact for h = 0 and h = 1, and without
    sin terms the generic sums add sin(m phi) x 0.  STRIDE1 / STRIDE2 are set
    exactly when every recurring group steps by 1 / 2, COS_ONLY exactly when
    no term has m < 0 (halo_lmax8: every recurring |m| group steps by 2 and
    holds cos terms only; the dense tables step by 1 and hold both)."""
    from agama_gpu_amd import MultipolePotentialGPU, _lib
    from agama_gpu_amd.multipole_prep import program_flags
    gpu = MultipolePotentialGPU(cases[name])
    prog = gpu._data["program"]
    # (halo_lmax8 is also the triaxial program of degree 8, bits 8-15, and the
    # dense tables the full program of degree 8, bits 16-23, whose unrolled
    # kernels tests/test_triaxial_gpu.py and tests/test_full_program_gpu.py
    # check; off here)
    want = (_lib.AGX_PROG_COS_ONLY | _lib.AGX_PROG_STRIDE2 | (8 << 8)) \
        if name == "halo_lmax8" else _lib.AGX_PROG_STRIDE1 | (8 << 16)
    assert program_flags(prog) == want == gpu._desc.prog_flags
    x = torch.as_tensor(points(100_000, 4), device="cuda")
    _lib.set_option(_lib.AGX_OPT_MP_PATH, 3)
    _lib.set_option(_lib.AGX_OPT_MP_TRIAXIAL, 0)
    try:
        outs = []
        for flags in (gpu._desc.prog_flags, 0):
            gpu._desc.prog_flags = flags
            outs.append(gpu.evalDeriv(x))
    finally:
        gpu._desc.prog_flags = want
        _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)
        _lib.set_option(_lib.AGX_OPT_MP_TRIAXIAL, 1)
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("name", ["fire_dense", "lmc", "halo_lmax8"])
def test_composite_program_flag_variants(cases, name):
    """The fused composite kernel picks its Multipole code from the AND of the
    leaves' program flags: with a table's own flags it gives bit-for-bit the
    result of the generic code (flags 0), and a composite of a stride-1 and a
    stride-2 table (AND = 0) equals the sum of the two tables' own results.
    The exact origin is left out: the NFW leaf's force is NaN at r = 0."""
    from agama_gpu_amd import MultipolePotentialGPU, PotentialGPU
    gpu = MultipolePotentialGPU(cases[name])
    nfw = PotentialGPU(type="NFW", mass=1e12, scaleRadius=20.0)
    x = torch.as_tensor(points(50_000, 7)[1:], device="cuda")
    own = gpu._desc.prog_flags
    assert own != 0
    outs = []
    try:
        for flags in (own, 0):
            gpu._desc.prog_flags = flags
            outs.append((gpu + nfw).evalDeriv(x))
    finally:
        gpu._desc.prog_flags = own
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    other = MultipolePotentialGPU(cases["fire_dense" if name == "halo_lmax8" else "halo_lmax8"])
    assert gpu._desc.prog_flags & other._desc.prog_flags == 0
    mixed = (gpu + other).evalDeriv(x)
    parts = [p.evalDeriv(x) for p in (gpu, other)]
    for k, m in enumerate(mixed):
        ref = parts[0][k] + parts[1][k]
        assert torch.max(torch.abs(m - ref)) <= 1e-13 * torch.max(torch.abs(ref))


@pytest.mark.parametrize("cos_only", [False, True])
def test_mixed_stride_program(cases, cos_only, mp_path):
    """A program whose |m| groups step by different strides: the dense table
    with its odd-l m = 0 harmonics zeroed (pruned), so the m = 0 group steps
    by 2 while the others step by 1; with cos_only also the sin harmonics
    zeroed.  No stride flag applies, so every path runs the per-group stride
    arithmetic; against the oracle on the same zeroed table."""
    from agama_gpu_amd import MultipolePotentialGPU, _lib
    from agama_gpu_amd.multipole_prep import program_flags
    from oracle.oracle import OracleMultipole

    def keep(l, m):
        return not (m == 0 and l % 2 == 1) and (m >= 0 or not cos_only)
    c = cases["fire_dense"].zeroed(keep)
    print(f"mixed strides cos_only={cos_only} path={mp_path}", flush=True)
    gpu = MultipolePotentialGPU(c)
    prog = gpu._data["program"]
    assert set(prog.grp_stride[np.diff(prog.grp_step0) > 1].tolist()) == {1, 2}
    assert program_flags(prog) == (_lib.AGX_PROG_COS_ONLY if cos_only else 0)
    orc = OracleMultipole.from_coefs(c)
    x = points(50_000, 3)
    xt = torch.as_tensor(x, device="cuda")
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(xt)]
    po, Fo, Do = orc.eval_deriv(x)
    assert rel_phi(p, po) < TOL_PHI
    assert rel_glob(F, Fo) < TOL_F
    R = np.hypot(x[:, 0], x[:, 1])
    ok = (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))
    assert rel_glob(D[ok], Do[ok]) < TOL_D
    p2, F2 = [t.cpu().numpy() for t in gpu.eval(xt, pot=True, acc=True)]
    assert rel_phi(p2, po) < TOL_PHI and rel_glob(F2, Fo) < TOL_F
