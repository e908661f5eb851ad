"""Wrappers: Composite, Evolving, Shifted and Scaled potentials.

Same semantics as the reference classes (gpu_potential.py: Composite
1439-1529, Evolving 1536-1644, Shifted 1651-1764, Scaled 1767-1918), but
each wrapper only rewrites the leaf list of what it wraps (see leaves.py), so
any nesting of them evaluates in one fused launch instead of a launch per
component plus elementwise glue kernels.  Time-dependent centres, scales and
amplitudes are splines evaluated on the host at the call's t, as in the
reference.
"""
from __future__ import annotations

from typing import List

import numpy as np

from .base import PotentialBase


class CompositePotentialGPU(PotentialBase):
    """Sum of GPU potential components (any mix of types)."""

    def __init__(self, components: list) -> None:
        if not components:
            raise ValueError("CompositePotentialGPU requires at least one component.")
        super().__init__(components[0].device)
        self._components = list(components)

    def __repr__(self):
        return f"CompositePotentialGPU({len(self._components)} components)"

    def __len__(self):
        return len(self._components)

    def __getitem__(self, i):
        return self._components[i]

    @property
    def components(self):
        return list(self._components)

    def _leaves(self, t=0.0):
        out = []
        for c in self._components:
            out.extend(c._leaves(t))
        return out


class EvolvingPotentialGPU(PotentialBase):
    """Snapshots at known times; linear interpolation in time
    (``interpolate=True``) or the nearest snapshot."""

    def __init__(self, potentials: list, times, interpolate: bool = True) -> None:
        if len(potentials) != len(times):
            raise ValueError(f"len(potentials)={len(potentials)} != len(times)={len(times)}")
        if len(potentials) < 1:
            raise ValueError("Need at least one snapshot.")
        super().__init__(potentials[0].device)
        self._pots = list(potentials)
        self._times = np.asarray(times, dtype=np.float64)
        self._interpolate = interpolate

    def _bracket(self, t: float):
        """(i, alpha) with t in [times[i], times[i+1]], clamped outside
        (gpu_potential.py:1559-1575)."""
        times, n = self._times, len(self._times)
        if t <= times[0] or n == 1:
            return 0, 0.0
        if t >= times[-1]:
            return n - 2, 1.0
        i = int(np.searchsorted(times, t, side="right") - 1)
        i = min(max(i, 0), n - 2)
        return i, float((t - times[i]) / (times[i + 1] - times[i]))

    def weights(self, t: float):
        """[(snapshot index, weight)] used at time t (gpu_potential.py:1577-1589)."""
        t = float(t)
        if len(self._pots) == 1 or not self._interpolate:
            return [(int(np.argmin(np.abs(self._times - t))), 1.0)]
        i, a = self._bracket(t)
        if a == 0.0:
            return [(i, 1.0)]
        if a == 1.0:
            return [(i + 1, 1.0)]
        return [(i, 1.0 - a), (i + 1, a)]

    def _leaves(self, t=0.0):
        out = []
        for i, w in self.weights(t):
            out.extend(lf.weighted(w) for lf in self._pots[i]._leaves(t))
        return out


def _linear_extrapolating_spline(times, values, derivs=None):
    """Cubic spline in t with linear extrapolation outside [t0, tN]: not-a-knot
    through the values, or Hermite when derivatives are given
    (gpu_potential.py:1683-1736, 1817-1846)."""
    from scipy.interpolate import CubicHermiteSpline, CubicSpline
    spl = CubicHermiteSpline(times, values, derivs) if derivs is not None else \
        CubicSpline(times, values, bc_type="not-a-knot")
    t0, tN = float(times[0]), float(times[-1])
    v0, vN, d0, dN = spl(t0), spl(tN), spl(t0, 1), spl(tN, 1)

    def f(t):
        t = float(t)
        if t < t0:
            return v0 + d0 * (t - t0)
        if t > tN:
            return vN + dN * (t - tN)
        return spl(t)
    return f


class ShiftedPotentialGPU(PotentialBase):
    """Evaluates ``inner`` at ``x - center(t)``; center is (3,) static, (T,4)
    rows [t, x, y, z] (cubic spline) or (T,7) rows [t, x, y, z, vx, vy, vz]
    (cubic Hermite), linearly extrapolated outside the time range."""

    def __init__(self, inner, center) -> None:
        super().__init__(inner.device)
        self._inner = inner
        c = np.asarray(center, dtype=np.float64)
        if c.ndim == 1 and c.shape == (3,):
            self._static = c.copy()
            self._fn = None
        elif c.ndim == 2 and c.shape[1] >= 4:
            if not np.all(c[1:, 0] > c[:-1, 0]):
                c = c[np.argsort(c[:, 0])]
            self._static = None
            self._times = c[:, 0].copy()
            self._fn = _linear_extrapolating_spline(
                c[:, 0], c[:, 1:4], c[:, 4:7] if c.shape[1] >= 7 else None)
        else:
            raise ValueError("Center must be (3,) or (T, 4) or (T, 7).")

    def center_at(self, t: float) -> np.ndarray:
        return self._static if self._fn is None else np.asarray(self._fn(t), dtype=np.float64)

    def _leaves(self, t=0.0):
        c = self.center_at(t)
        return [lf.shifted(c) for lf in self._inner._leaves(t)]


class ScaledPotentialGPU(PotentialBase):
    """Phi_s(x, t) = ampl(t) * s(t) * Phi(x * s(t)), s = 1/scale(t); scale is a
    float, (T,2) rows [t, scale] or (T,3) rows [t, ampl, scale]
    (gpu_potential.py:1767-1918)."""

    def __init__(self, inner, scale, ampl: float = 1.0) -> None:
        super().__init__(inner.device)
        self._inner = inner
        arr = None if isinstance(scale, (int, float)) else np.asarray(scale, dtype=np.float64)
        if arr is not None and arr.ndim == 0:
            arr = None
        self._ampl_val = float(ampl)
        if arr is None:
            self._scale_fn = None
            self._scale_val = float(scale)
            self._ampl_fn = None
            return
        if arr.ndim != 2 or arr.shape[1] not in (2, 3):
            raise ValueError("scale must be a float, (T,2) array [t, scale(t)], or (T,3) "
                             f"array [t, ampl(t), scale(t)]. Got shape {arr.shape}.")
        if not np.all(arr[1:, 0] > arr[:-1, 0]):
            arr = arr[np.argsort(arr[:, 0])]
        self._scale_fn = _linear_extrapolating_spline(arr[:, 0], arr[:, -1])
        self._ampl_fn = _linear_extrapolating_spline(arr[:, 0], arr[:, 1]) if arr.shape[1] == 3 else None

    def s_and_a(self, t: float):
        """(1/scale(t), ampl(t)) (gpu_potential.py:1848-1876)."""
        sc = self._scale_val if self._scale_fn is None else float(self._scale_fn(t))
        a = self._ampl_val if self._ampl_fn is None else float(self._ampl_fn(t))
        return 1.0 / sc, a

    def _leaves(self, t=0.0):
        s, a = self.s_and_a(t)
        return [lf.scaled(s, a) for lf in self._inner._leaves(t)]
