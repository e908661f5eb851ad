"""The unrolled LDS-staged kernels of triaxial programs (multipole_tri.hip).

A table holding every harmonic that triaxial symmetry allows up to an even
degree L (l even, m even and >= 0) has a fixed program (triaxial_degree);
the host names L in prog_flags bits 8-15 and the LDS path runs a kernel with
that program unrolled at compile time (degrees 2..12, the potential alone
2..10; tables above 80 KB run the generic kernels).  Checked here, on the
LDS path (N >= 4096):

* for triaxial Dehnen tables of degree 2..12, every evaluation order against
  the generic kernel of the same table (AGX_OPT_MP_TRIAXIAL 0) and against
  the oracle, at points over both extrapolation regions, on the z axis, at
  the origin and at r_max;
* a 99-interval degree-8 table (71 KB: the 512-thread variant);
* a triaxial table with one harmonic pruned: no degree, the generic kernel.
Tolerances: against the generic kernel 1e-14 of the maximum (the unrolled
code contracts some products into FMAs differently); against the oracle those
of the fixture tests: phi 1e-12, force 1e-11, deriv and density 1e-10
(rows with 0 < sin(theta) < 1e-6 excluded from deriv, as there).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEGREES = [2, 4, 6, 8, 10, 12]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def regular(x):
    R = np.hypot(x[:, 0], x[:, 1])
    return (R == 0) | (R > 1e-6 * np.linalg.norm(x, axis=1))


def sample(n, r_max, seed):
    from agama_gpu_amd.workloads import hernquist_points
    x = hernquist_points(n, seed)
    x[:8] = [[0, 0, 0], [0, 0, 3.0], [0, 0, -3.0], [1e-12, 0, 7.0], [0, 0, 1e-3], [5.0, -2.0, 0],
             [0, 0, r_max * (1 - 1e-9)], [r_max * (1 + 1e-9), 0, 0]]
    return x


@pytest.fixture(scope="module")
def tables():
    from agama_gpu_amd.workloads import dehnen_density, multipole_from_density
    rho = dehnen_density(1e12, 20.0, 1.0, 0.8, 0.6)
    out = {L: multipole_from_density(rho, L, 0.1, 500.0, 50, symmetry="triaxial") for L in DEGREES}
    out["n100"] = multipole_from_density(rho, 8, 0.1, 500.0, 100, symmetry="triaxial")
    return out


@pytest.fixture()
def lds_path():
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_MP_PATH, 3)
    yield
    _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)
    _lib.set_option(_lib.AGX_OPT_MP_TRIAXIAL, 1)


def evaluate(gpu, x, tri):
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_MP_TRIAXIAL, tri)
    xt = torch.as_tensor(x, device="cuda")
    out = [gpu.potential(xt)]
    out += list(gpu.eval(xt, pot=True, acc=True))
    out += list(gpu.evalDeriv(xt))
    out.append(gpu.density(xt))
    return [t.cpu().numpy() for t in out]


def check(gpu, c, x, what):
    from oracle.oracle import OracleMultipole
    tri = evaluate(gpu, x, 1)
    gen = evaluate(gpu, x, 0)
    for a, b in zip(tri, gen):
        assert rel(a, b) <= 1e-14, what
    o = OracleMultipole.from_coefs(c)
    po, Fo, Do = o.eval_deriv(x)
    p, p2, F2, p3, F3, D3, rho = tri
    for pv in (p, p2, p3):
        assert rel(pv, po) < 1e-12, what
    for Fv in (F2, F3):
        assert rel(Fv, Fo) < 1e-11, what
    ok = regular(x)
    assert rel(D3[ok], Do[ok]) < 1e-10, what
    assert np.all(np.isfinite(D3)), what
    assert rel(rho, o.density(x)) < 1e-10, what


@pytest.mark.parametrize("L", DEGREES)
def test_triaxial_degree_against_generic_and_oracle(tables, lds_path, L):
    from agama_gpu_amd import MultipolePotentialGPU, _lib
    c = tables[L]
    gpu = MultipolePotentialGPU(c)
    degree = (gpu._desc.prog_flags & _lib.AGX_PROG_TRIAXIAL_MASK) >> _lib.AGX_PROG_TRIAXIAL_SHIFT
    print(f"case triaxial L={L} n_lm={gpu._desc.n_lm} flags={gpu._desc.prog_flags}", flush=True)
    assert degree == L and gpu._desc.n_lm == (L // 2 + 1) * (L // 2 + 2) // 2
    check(gpu, c, sample(20_000, float(c.R_grid[-1]), L), f"L={L}")


def test_triaxial_512_thread_variant(tables, lds_path):
    """99 radial intervals of 15 harmonics: 71 KB, two workgroups of 512
    threads per CU."""
    from agama_gpu_amd import MultipolePotentialGPU
    c = tables["n100"]
    gpu = MultipolePotentialGPU(c)
    assert (gpu._desc.prog_flags >> 8) & 0xFF == 8
    assert 40 * 1024 < gpu._desc.n_int * (3 * gpu._desc.n_lm) * 16 <= 80 * 1024
    print("case triaxial L=8 nR=100", flush=True)
    check(gpu, c, sample(30_000, float(c.R_grid[-1]), 5), "nR=100")


def test_triaxial_bad_positions(tables, lds_path):
    """NaN in any coordinate gives NaN outputs, r^2 = inf gives 0 (the outer
    law at r = inf), as on the other kernels; the good points around them
    are those of the same launch without the bad ones, bit for bit."""
    from agama_gpu_amd import MultipolePotentialGPU, _lib
    gpu = MultipolePotentialGPU(tables[8])
    _lib.set_option(_lib.AGX_OPT_MP_TRIAXIAL, 1)
    good = sample(8192, float(tables[8].R_grid[-1]), 11)
    bad = np.array([[np.nan, 1.0, 2.0], [1.0, np.nan, 2.0], [1.0, 2.0, np.nan],
                    [1e200, 0.0, 0.0], [0.0, -1e200, 3.0], [1.0, 1.0, 1e170]])
    isnan = np.array([True, True, True, False, False, False])
    print("case triaxial L=8 bad positions", flush=True)
    x = torch.as_tensor(np.concatenate([good, bad, good[:100]]), device="cuda")
    p, F, D = [t.cpu().numpy() for t in gpu.evalDeriv(x)]
    rp, rF, rD = [t.cpu().numpy() for t in gpu.evalDeriv(torch.as_tensor(good, device="cuda"))]
    b = slice(8192, 8192 + len(bad))
    assert np.all(np.isnan(p[b][isnan])) and np.all(np.isnan(F[b][isnan]).any(1))
    assert np.all(np.isnan(D[b][isnan]).any(1))
    assert np.all(p[b][~isnan] == 0.0) and np.all(F[b][~isnan] == 0.0)
    assert np.all(D[b][~isnan] == 0.0)
    assert np.array_equal(p[:8192], rp) and np.array_equal(F[:8192], rF)
    assert np.array_equal(D[:8192], rD)


def test_pruned_triaxial_table_runs_generic(tables, lds_path):
    """One harmonic zeroed (pruned from the program): no triaxial degree, and
    the option changes nothing."""
    from agama_gpu_amd import MultipolePotentialGPU
    c = tables[8].zeroed(lambda l, m: (l, m) != (6, 4))
    gpu = MultipolePotentialGPU(c)
    assert gpu._desc.prog_flags & 0xFF00 == 0
    x = sample(10_000, float(c.R_grid[-1]), 9)
    for a, b in zip(evaluate(gpu, x, 1), evaluate(gpu, x, 0)):
        assert np.array_equal(a, b)
