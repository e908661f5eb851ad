"""CompositePotentialGPU -- sum of components (gpu_potential.py:1439-1529)."""
from __future__ import annotations

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

    @property
    def components(self):
        return list(self._components)

    def _evaluate(self, xyz, want_phi, want_force, want_deriv, t=0.0, out=None):
        acc = None
        for i, c in enumerate(self._components):
            r = c._evaluate(xyz, want_phi, want_force, want_deriv, t, out=out if i == 0 else None)
            if acc is None:
                acc = list(r)
            else:
                for i in range(3):
                    if r[i] is not None:
                        acc[i].add_(r[i])
        return tuple(acc)

    def _density(self, xyz, t=0.0):
        rho = None
        for c in self._components:
            d = c._density(xyz, t)
            rho = d if rho is None else rho.add_(d)
        return rho
