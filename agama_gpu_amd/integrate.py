"""Leapfrog integration of test particles through GPU potentials.

``LeapfrogGPU(potential, dt, n_steps, t0)`` tabulates the potential's time
dependence once on the host -- for every step s, the flattened leaves
(Evolving weights, Shifted centres, Scaled factors) at t0 + (s+1) dt -- and
uploads it as a schedule (agx_leapfrog in include/agama_gpu.h).  A step is
one kernel launch (kick, drift, force at the new time, kick) plus a one-thread
launch that advances the device-resident step counter, so a captured HIP
graph of K steps replays without any host work; ``run(graph=True)`` captures
one such graph (torch.cuda.CUDAGraph) and replays it.

``run_eager`` is the same integration written with the public evaluation API
(one force evaluation and four elementwise updates per step, time passed from
the host): the per-step launch sequence the graph replaces, and the check of
its result.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Tuple

import numpy as np
import torch

from . import _lib
from .base import prep_xyz

F64 = torch.float64


class LeapfrogGPU:
    """Kick-drift-kick leapfrog with a fixed step through ``potential``."""

    def __init__(self, potential, dt: float, n_steps: int, t0: float = 0.0):
        if n_steps < 1:
            raise ValueError("n_steps must be >= 1")
        self.potential = potential
        self.device = potential.device
        self.dt, self.n_steps, self.t0 = float(dt), int(n_steps), float(t0)
        keys: Dict[Tuple, int] = {}
        leaves: List = []
        rows: List[List[Tuple[int, float, float, Tuple[float, float, float]]]] = []
        for s in range(self.n_steps):
            row = []
            for lf in potential._leaves(self.time_of_step(s + 1)):
                if lf.ampl == 0.0:
                    continue
                key = (lf.type, tuple(float(p) for p in lf.params), int(lf.table or 0))
                if key not in keys:
                    keys[key] = len(leaves)
                    leaves.append(lf)
                row.append((keys[key], lf.ampl, lf.sigma, lf.tau))
            rows.append(row)
        self.max_active = max(1, max(len(r) for r in rows))
        sched = np.zeros((self.n_steps, self.max_active), _lib.SCHED_ENTRY_DTYPE)
        sched["leaf"] = -1
        count = np.zeros(self.n_steps, np.int32)
        for s, row in enumerate(rows):
            count[s] = len(row)
            for e, (li, a, sg, tau) in enumerate(row):
                sched[s, e] = (li, 0, a, sg, tau)
        from .leaves import to_components
        comps = to_components(leaves) if leaves else (_lib.agx_component * 1)()
        self.leaves = leaves                    # the distinct leaves, in schedule numbering
        self.n_leaves = len(leaves)
        dev = self.device
        self._leaves = torch.as_tensor(
            np.frombuffer(bytes(comps), np.uint8).copy(), device=dev)
        self._sched = torch.as_tensor(sched.view(np.uint8).ravel().copy(), device=dev)
        self._count = torch.as_tensor(count, device=dev)
        self._step = torch.zeros(1, dtype=torch.int32, device=dev)
        self._graph = None
        self._graph_k = 0
        self.x = self.v = self.a = None

    def time_of_step(self, s: int) -> float:
        return self.t0 + s * self.dt

    # ---- state ---------------------------------------------------------------
    def set_state(self, x, v):
        """Copy positions and velocities in, rewind to t0 and evaluate the
        initial acceleration."""
        x, _ = prep_xyz(x, self.device)
        v, _ = prep_xyz(v, self.device)
        if x.shape != v.shape:
            raise ValueError("x and v must have the same shape")
        if self.x is None or self.x.shape != x.shape:
            self.x = torch.empty_like(x)
            self.v = torch.empty_like(v)
            self.a = torch.empty_like(x)
            self._graph = None
        self.x.copy_(x)
        self.v.copy_(v)
        self.potential.eval_into(self.x, force=self.a, t=self.t0)
        self._step.zero_()

    @property
    def steps_done(self) -> int:
        return int(self._step.item())

    def _launch(self, k: int):
        rc = _lib.load().agx_leapfrog(
            self._leaves.data_ptr(), self.n_leaves, self._sched.data_ptr(), self._count.data_ptr(),
            self.max_active, self.n_steps, self._step.data_ptr(), self.x.data_ptr(),
            self.v.data_ptr(), self.a.data_ptr(), self.x.shape[0], self.dt, k,
            _lib.stream_handle(self.device))
        _lib.check(rc, "agx_leapfrog")

    def run(self, x, v, n_steps: int = None, graph: bool = True, steps_per_graph: int = 50):
        """Integrate n_steps (default: all scheduled) from (x, v) at t0;
        returns the final (x, v) tensors (owned by the integrator)."""
        n = self.n_steps if n_steps is None else int(n_steps)
        if n > self.n_steps:
            raise ValueError(f"only {self.n_steps} steps are scheduled")
        self.set_state(x, v)
        self.advance(n, graph=graph, steps_per_graph=steps_per_graph)
        return self.x, self.v

    def advance(self, n: int, graph: bool = True, steps_per_graph: int = 50):
        """n more steps from the current state."""
        if not graph:
            self._launch(n)
            return
        k = max(1, min(int(steps_per_graph), n))
        if self._graph is None or self._graph_k != k:
            self._graph = torch.cuda.CUDAGraph()
            self._graph_k = k
            # capture launches do not execute; the step counter is untouched
            with torch.cuda.graph(self._graph):
                self._launch(k)
        full, rest = divmod(n, k)
        for _ in range(full):
            self._graph.replay()
        if rest:
            self._launch(rest)

    # ---- reference form --------------------------------------------------------
    def run_eager(self, x, v, n_steps: int = None):
        """The same integration through the public API, one step at a time."""
        n = self.n_steps if n_steps is None else int(n_steps)
        x, _ = prep_xyz(x, self.device)
        v, _ = prep_xyz(v, self.device)
        x, v = x.clone(), v.clone()
        a = self.potential.force(x, t=self.t0)
        h = 0.5 * self.dt
        for s in range(n):
            v.add_(a, alpha=h)
            x.add_(v, alpha=self.dt)
            a = self.potential.force(x, t=self.time_of_step(s + 1))
            v.add_(a, alpha=h)
        return x, v
