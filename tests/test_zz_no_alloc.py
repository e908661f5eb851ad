"""agx_multipole_eval_sorted_noalloc (the interval-sorted Multipole path with
its sort in the library's static device scratch) against
agx_multipole_eval_sorted with a caller workspace, through the C ABI: the
same kernels on the same points, so the outputs agree bit for bit.  Two
calls in a row, the second on fewer points (the scratch keeps the first
call's sort; the second must not read it), for phi + force + deriv and for
pot+force, on the dense FIRE table and the lmax-32 table."""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import FIXTURES

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _outputs(n):
    return (torch.full((n,), np.nan, dtype=torch.float64, device=DEV),
            torch.full((n, 3), np.nan, dtype=torch.float64, device=DEV),
            torch.full((n, 6), np.nan, dtype=torch.float64, device=DEV))


@pytest.mark.parametrize("table", ["fire_dense", "lmax32"])
def test_noalloc_matches_workspace_entry(table):
    from agama_gpu_amd import MultipolePotentialGPU, _lib, read_coefs
    from agama_gpu_amd.workloads import high_order_coefs, hernquist_points, loguniform_points
    L = _lib.load()
    if table == "fire_dense":
        c = read_coefs(FIXTURES / "600.dark.none_8.coef_mul_DR")
        x_all = hernquist_points(300_000, 21)
    else:
        c = high_order_coefs()
        x_all = loguniform_points(300_000, 22, 0.05, 800.0)
    pot = MultipolePotentialGPU(c, device=DEV)
    desc = C.byref(pot._desc)
    stream = _lib.stream_handle(torch.device(DEV))
    full = _lib.AGX_PHI | _lib.AGX_FORCE | _lib.AGX_DERIV
    for n, what in ((300_000, full), (50_001, full), (70_000, _lib.AGX_PHI | _lib.AGX_FORCE)):
        print(f"noalloc table={table} n={n} what={what}", flush=True)
        xyz = torch.as_tensor(x_all[:n], device=DEV)
        a = _outputs(n)
        b = _outputs(n)
        nwork = int(L.agx_multipole_work_size(n))
        work = torch.empty(nwork, dtype=torch.uint8, device=DEV)
        rc = L.agx_multipole_eval_sorted(desc, xyz.data_ptr(), n, a[0].data_ptr(), a[1].data_ptr(),
                                         a[2].data_ptr(), what, work.data_ptr(), nwork, stream)
        assert rc == 0
        rc = L.agx_multipole_eval_sorted_noalloc(desc, xyz.data_ptr(), n, b[0].data_ptr(),
                                                 b[1].data_ptr(), b[2].data_ptr(), what, stream)
        assert rc == 0
        torch.cuda.synchronize()
        k = 3 if what & _lib.AGX_DERIV else 2
        for u, v in zip(a[:k], b[:k]):
            assert torch.isfinite(v).all()
            assert torch.equal(u, v)
        if k == 2:                            # deriv not requested: left untouched
            assert torch.isnan(b[2]).all()
    # beyond the static scratch: rejected on the host (NULL pointers, so a
    # check that regressed could not launch either)
    big = _lib.AGX_NOALLOC_MAX_N + 1
    assert L.agx_multipole_eval_sorted_noalloc(desc, None, big, None, None, None, full,
                                               stream) == -8
