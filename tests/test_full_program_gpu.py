"""The unrolled LDS-staged kernels of full programs (multipole_full.hip).

A table holding every harmonic up to a degree L, cos and sin (no symmetry,
nothing pruned: the reference's own FIRE and LMC fixtures, lmax 8 on 25
nodes, and Agama's files of that kind generally) has a fixed program
(full_degree); the host names L in prog_flags bits 16-23 and the LDS path
runs a kernel with that program unrolled at compile time (even degrees
2..10).  The interval-sorted kernel's LDS-slab variant runs the same unrolled programs.
Checked here, on the LDS path and on the sorted path, every evaluation order:

* the two fixtures and random self-consistent full tables of degree 2..10
  (tests/random_tables.py, 19..26 nodes so that they fit in LDS) against the
  generic kernel of the same table (AGX_OPT_MP_TRIAXIAL 0) and the oracle, at
  points over both extrapolation regions, on the z axis, at the origin and at
  r_max;
* an odd degree (7: no unrolled variant) and a full table with one harmonic
  pruned (no degree): the generic kernel, bit for bit with the option off.
Tolerances as in tests/test_triaxial_gpu.py: 1e-14 of the maximum against the
generic kernel; phi 1e-12, force 1e-11, deriv and density 1e-10 against the
oracle (rows with 0 < sin(theta) < 1e-6 excluded from deriv).
"""
import numpy as np
import pytest
import torch

from random_tables import random_multipole
from test_triaxial_gpu import check, evaluate, lds_path, sample  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[3, 4], ids=["lds", "sorted"])
def lds_or_sorted(request):
    from agama_gpu_amd import _lib
    _lib.set_option(_lib.AGX_OPT_MP_PATH, request.param)
    yield request.param
    _lib.set_option(_lib.AGX_OPT_MP_PATH, 0)
    _lib.set_option(_lib.AGX_OPT_MP_TRIAXIAL, 1)

FIXTURES = {"fire": "600.dark.none_8.coef_mul_DR", "lmc": "100.LMC.none_8.coef_mult"}


def full_table(L, seed0):
    """The first random full table of degree L from seed0 on with 20..26
    nodes (LDS-sized)."""
    for s in range(seed0, seed0 + 400):
        c, d = random_multipole(s, lmax=L, pattern="full", signed=s % 2 == 1)
        if 20 <= d["nR"] <= 26:
            return c, d
    raise RuntimeError("no table")


@pytest.mark.parametrize("name", list(FIXTURES))
def test_full_program_fixtures(lds_or_sorted, name):
    from pathlib import Path

    from agama_gpu_amd import MultipolePotentialGPU, _lib
    from agama_gpu_amd.coefs import read_coefs
    c = read_coefs(Path(__file__).parent / "golden" / FIXTURES[name])
    gpu = MultipolePotentialGPU(c)
    degree = (gpu._desc.prog_flags & _lib.AGX_PROG_FULL_MASK) >> _lib.AGX_PROG_FULL_SHIFT
    print(f"case full fixture {name} flags={gpu._desc.prog_flags:#x} path={lds_or_sorted}",
          flush=True)
    assert degree == 8 and gpu._desc.n_lm == 81
    check(gpu, c, sample(20_000, float(c.R_grid[-1]), 3), name)


@pytest.mark.parametrize("L", [2, 4, 6, 8, 10])
def test_full_program_random_tables(lds_or_sorted, L):
    from agama_gpu_amd import MultipolePotentialGPU, _lib
    c, d = full_table(L, 100 * L)
    gpu = MultipolePotentialGPU(c)
    degree = (gpu._desc.prog_flags & _lib.AGX_PROG_FULL_MASK) >> _lib.AGX_PROG_FULL_SHIFT
    print(f"case full L={L} path={lds_or_sorted} {d}", flush=True)
    assert degree == L and gpu._desc.n_lm == (L + 1) ** 2
    check(gpu, c, sample(20_000, d["r_max"], L), d)


@pytest.mark.parametrize("kind", ["odd_degree", "pruned"])
def test_full_program_fallbacks(lds_path, kind):
    from agama_gpu_amd import MultipolePotentialGPU
    if kind == "odd_degree":
        c, d = full_table(7, 700)
        want = 7 << 16
    else:
        c, d = full_table(6, 600)
        c = c.zeroed(lambda l, m: (l, m) != (5, -3))
        want = 0
    gpu = MultipolePotentialGPU(c)
    assert gpu._desc.prog_flags & 0xFF0000 == want
    x = sample(10_000, d["r_max"], 9)
    for a, b in zip(evaluate(gpu, x, 1), evaluate(gpu, x, 0)):
        assert np.array_equal(a, b)
