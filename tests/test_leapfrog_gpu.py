"""Leapfrog through time-dependent potentials on gfx950 (agx_leapfrog).

* Graph replay (device-resident step counter) equals the step-by-step
  integration through the public API with the time passed from the host,
  for an Evolving Multipole halo plus a spline-shifted satellite (30 steps,
  graphs of 7 steps plus a remainder): 1e-11 of the position/velocity scale.
* Resuming (advance after run) continues the same trajectory.
* Leapfrog is symplectic: in a static NFW potential the energy error stays
  bounded (below 1e-5 relative over 400 steps) and the integration is
  time-reversible (flip v, integrate back: initial state to 1e-9).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(scope="module")
def problem():
    from agama_gpu_amd.workloads import leapfrog_problem
    pot, x, v = leapfrog_problem(3000, device=DEV, n_snap=4)
    return pot, x, v


def _rel(a, b):
    return (torch.max(torch.abs(a - b)) / torch.max(torch.abs(b))).item()


def test_graph_equals_eager(problem):
    from agama_gpu_amd import LeapfrogGPU
    pot, x, v = problem
    print("case evolving+satellite n=3000 steps=30", flush=True)
    lf = LeapfrogGPU(pot, dt=2e-3, n_steps=30, t0=0.25)
    xg, vg = lf.run(x, v, graph=True, steps_per_graph=7)
    torch.cuda.synchronize()
    assert lf.steps_done == 30
    xe, ve = lf.run_eager(x, v)
    assert _rel(xg, xe) < 1e-11
    assert _rel(vg, ve) < 1e-11
    # without a graph: the same launches, eagerly
    xn, vn = lf.run(x, v, graph=False)
    assert torch.equal(xn, xg) and torch.equal(vn, vg)


def test_resume(problem):
    from agama_gpu_amd import LeapfrogGPU
    pot, x, v = problem
    lf = LeapfrogGPU(pot, dt=2e-3, n_steps=24, t0=0.0)
    lf.run(x, v, n_steps=10, graph=True, steps_per_graph=5)
    lf.advance(14, graph=True, steps_per_graph=5)
    assert lf.steps_done == 24
    xe, ve = lf.run_eager(x, v, 24)
    assert _rel(lf.x, xe) < 1e-11 and _rel(lf.v, ve) < 1e-11


def test_steps_past_schedule_leave_particles(problem):
    """A call that runs past the schedule (step counter + j >= n_sched) moves
    the particles only for its scheduled steps; the counter still advances by
    the call's step count."""
    from agama_gpu_amd import LeapfrogGPU
    pot, x, v = problem
    lf = LeapfrogGPU(pot, dt=2e-3, n_steps=12, t0=0.0)
    lf.run(x, v, n_steps=9, graph=False)
    lf.advance(6, graph=False)                     # steps 9..11 run, 12..14 do nothing
    assert lf.steps_done == 15
    xe, ve = lf.run_eager(x, v, 12)
    assert _rel(lf.x, xe) < 1e-11 and _rel(lf.v, ve) < 1e-11
    x_end = lf.x.clone()
    lf.advance(4, graph=True, steps_per_graph=4)
    assert lf.steps_done == 19 and torch.equal(lf.x, x_end)


def _energy(pot, x, v):
    return pot.potential(x) + 0.5 * (v * v).sum(1)


def test_static_energy_and_reversibility():
    from agama_gpu_amd import LeapfrogGPU, NFWPotentialGPU
    from agama_gpu_amd.workloads import hernquist_points
    pot = NFWPotentialGPU(mass=1e12, scaleRadius=20.0, device=DEV)
    x = torch.as_tensor(hernquist_points(2000, 9), device=DEV)
    x = x[torch.sqrt((x * x).sum(1)) > 2.0].contiguous()
    # near-circular velocities (perpendicular to r in the x-y plane direction)
    r = torch.sqrt((x * x).sum(1, keepdim=True))
    vc = torch.sqrt(-(pot.force(x) * x).sum(1, keepdim=True))
    ez = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64, device=DEV).expand_as(x)
    t = torch.linalg.cross(ez, x)
    t = t / torch.clamp(torch.sqrt((t * t).sum(1, keepdim=True)), min=1e-30)
    v = 0.8 * vc * t
    lf = LeapfrogGPU(pot, dt=1e-4, n_steps=400)
    E0 = _energy(pot, x, v)
    xf, vf = lf.run(x, v, graph=True, steps_per_graph=50)
    E1 = _energy(pot, xf, vf)
    assert torch.max(torch.abs((E1 - E0) / E0)).item() < 1e-5
    back = LeapfrogGPU(pot, dt=1e-4, n_steps=400)
    xb, vb = back.run(xf.clone(), -vf.clone(), graph=True, steps_per_graph=50)
    assert _rel(xb, x) < 1e-9
    assert _rel(-vb, v) < 1e-9
