"""Host side of the leapfrog (LeapfrogGPU's schedule): leaves are shared
between steps, each step lists exactly the leaves of the potential at
t0 + (s+1) dt with their weights, scales and centres.  CPU only (tables on
the CPU device; nothing is launched)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def problem():
    from agama_gpu_amd import (EvolvingPotentialGPU, MultipolePotentialGPU, ShiftedPotentialGPU,
                               PlummerPotentialGPU)
    from agama_gpu_amd.workloads import halo_lmax8_coefs, satellite_track
    c = halo_lmax8_coefs()
    snaps = [MultipolePotentialGPU(c, device="cpu") for _ in range(4)]
    evo = EvolvingPotentialGPU(snaps, [0.0, 0.1, 0.2, 0.3])
    sat = ShiftedPotentialGPU(PlummerPotentialGPU(mass=1e10, scaleRadius=1.0, device="cpu"),
                              satellite_track(12, 0.5))
    return evo + sat, evo, sat


def test_schedule_rows_match_leaves(problem):
    from agama_gpu_amd import LeapfrogGPU
    from agama_gpu_amd import _lib
    pot, evo, sat = problem
    dt, n = 0.013, 31
    lf = LeapfrogGPU(pot, dt, n, t0=-0.02)
    sched = lf._sched.numpy().view(_lib.SCHED_ENTRY_DTYPE).reshape(n, lf.max_active)
    count = lf._count.numpy()
    assert lf.n_leaves == 5                       # 4 snapshots + 1 satellite, shared by all steps
    assert lf.max_active == 3
    for s in range(n):
        t = -0.02 + (s + 1) * dt
        want = [(i, w) for i, w in evo.weights(t)]
        row = sched[s, :count[s]]
        assert count[s] == len(want) + 1
        snap_of = {s_._dev_desc.data_ptr(): i for i, s_ in enumerate(evo._pots)}
        got = sorted((snap_of[lf.leaves[int(e["leaf"])].table], float(e["ampl"])) for e in row[:-1])
        assert got == sorted((int(i), float(w)) for i, w in want)
        c = sat.center_at(t)
        assert np.allclose(row[-1]["center"], c, rtol=0, atol=1e-12)
        assert row[-1]["scale"] == 1.0
    # padding entries name no leaf
    for s in range(n):
        assert np.all(sched[s, count[s]:]["leaf"] == -1)
