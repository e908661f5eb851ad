"""Analytic potentials (fused elementwise gfx950 kernels).

Same classes, constructor keywords and formulas as the reference's
``_analytic_potentials.py`` (NFW :176-281, Plummer :288-377, Hernquist
:383-480, spherical Dehnen :486-597, Isochrone :603-713, MiyamotoNagai
:720-835, LogHalo :842-931, DiskAnsatz :950-1106, UniformAcceleration
:1112-1165, AnalyticPotentialGPU factory :1173-1217).  Each class is one leaf
of the fused composite kernel (agx_composite_eval): a single launch returns
every requested output from one read of the positions.
"""
from __future__ import annotations

import math

from . import _lib
from .base import PotentialBase
from .leaves import Leaf
from .multipole_prep import G_AGAMA


class _Analytic(PotentialBase):
    _type = 0

    def __init__(self, device=None):
        super().__init__(device)

    def _params(self):
        raise NotImplementedError

    def _leaves(self, t=0.0):
        return [Leaf(self._type, tuple(self._params()))]

    @classmethod
    def from_agama(cls, pot):
        raise TypeError(
            f"{cls.__name__}.from_agama() is not supported: Agama does not export analytic "
            "potential parameters.  Construct the class directly.")

    def __repr__(self):
        return f"{type(self).__name__}({', '.join(f'{v:g}' for v in self._params())})"


class NFWPotentialGPU(_Analytic):
    """Phi = -G M ln(1 + r/rs) / r."""
    _type = _lib.AGX_T_NFW

    def __init__(self, mass: float = 1.0, scaleRadius: float = 1.0, device=None):
        super().__init__(device)
        self._M, self._rs = float(mass), float(scaleRadius)

    def _params(self):
        return (G_AGAMA * self._M, self._rs, self._M)


class PlummerPotentialGPU(_Analytic):
    """Phi = -G M / sqrt(r^2 + b^2)."""
    _type = _lib.AGX_T_PLUMMER

    def __init__(self, mass: float = 1.0, scaleRadius: float = 1.0, device=None):
        super().__init__(device)
        self._M, self._b = float(mass), float(scaleRadius)

    def _params(self):
        return (G_AGAMA * self._M, self._b, self._M)


class HernquistPotentialGPU(_Analytic):
    """Phi = -G M / (r + a)."""
    _type = _lib.AGX_T_HERNQUIST

    def __init__(self, mass: float = 1.0, scaleRadius: float = 1.0, device=None):
        super().__init__(device)
        self._M, self._a = float(mass), float(scaleRadius)

    def _params(self):
        return (G_AGAMA * self._M, self._a, self._M)


class DehnenSphericalPotentialGPU(_Analytic):
    """Phi = -(G M/a) (1 - (r/(r+a))^(2-gamma)) / (2-gamma); gamma = 2 gives
    -(G M/a) ln(1 + a/r)."""
    _type = _lib.AGX_T_DEHNEN

    def __init__(self, mass: float = 1.0, scaleRadius: float = 1.0, gamma: float = 1.0,
                 device=None):
        if not 0.0 <= gamma <= 2.0:
            raise ValueError(f"gamma must be in [0, 2], got {gamma}")
        super().__init__(device)
        self._M, self._a, self._g = float(mass), float(scaleRadius), float(gamma)

    def _params(self):
        return (G_AGAMA * self._M, self._a, self._g, self._M)


class IsochronePotentialGPU(_Analytic):
    """Phi = -G M / (b + sqrt(r^2 + b^2))."""
    _type = _lib.AGX_T_ISOCHRONE

    def __init__(self, mass: float = 1.0, scaleRadius: float = 1.0, device=None):
        super().__init__(device)
        self._M, self._b = float(mass), float(scaleRadius)

    def _params(self):
        return (G_AGAMA * self._M, self._b, self._M)


class MiyamotoNagaiPotentialGPU(_Analytic):
    """Phi = -G M / sqrt(R^2 + (a + sqrt(z^2 + b^2))^2)."""
    _type = _lib.AGX_T_MIYAMOTO_NAGAI

    def __init__(self, mass: float = 1.0, scaleRadius: float = 1.0, scaleHeight: float = 0.1,
                 device=None):
        super().__init__(device)
        self._M, self._a, self._b = float(mass), float(scaleRadius), float(scaleHeight)

    def _params(self):
        return (G_AGAMA * self._M, self._a, self._b, self._M)


class LogHaloPotentialGPU(_Analytic):
    """Phi = (v0^2/2) ln(rc^2 + x^2 + y^2/p^2 + z^2/q^2)."""
    _type = _lib.AGX_T_LOGHALO

    def __init__(self, velocity: float = 1.0, coreRadius: float = 0.01, axisRatioY: float = 1.0,
                 axisRatioZ: float = 1.0, device=None):
        super().__init__(device)
        self._v, self._rc, self._p, self._q = (float(velocity), float(coreRadius),
                                               float(axisRatioY), float(axisRatioZ))

    def _params(self):
        return (self._v ** 2, self._rc ** 2, self._p ** 2, self._q ** 2)


class DiskAnsatzPotentialGPU(_Analytic):
    """Phi = f(r) H(z), f = 4 pi G Sigma exp(-r/R_d - r_in/r); H exponential
    (scaleHeight > 0), isothermal sech^2 (< 0) or thin (= 0)."""
    _type = _lib.AGX_T_DISK_ANSATZ

    def __init__(self, surfaceDensity: float = 1.0, scaleRadius: float = 1.0,
                 scaleHeight: float = 0.1, innerCutoffRadius: float = 0.0, device=None):
        super().__init__(device)
        self._S, self._Rd, self._h, self._rin = (float(surfaceDensity), float(scaleRadius),
                                                 float(scaleHeight), float(innerCutoffRadius))
        if abs(self._h) < 1e-10:
            self._h = 0.0

    def _params(self):
        return (4.0 * math.pi * G_AGAMA * self._S, self._Rd, self._h, self._rin)


class UniformAccelerationGPU(_Analytic):
    """Phi = -(ax x + ay y + az z): a spatially constant force (ax, ay, az)."""
    _type = _lib.AGX_T_UNIFORM_ACCEL

    def __init__(self, ax: float = 0.0, ay: float = 0.0, az: float = 0.0, device=None):
        super().__init__(device)
        self._a = (float(ax), float(ay), float(az))

    def _params(self):
        return self._a


ANALYTIC_TYPES = {
    "nfw": NFWPotentialGPU,
    "plummer": PlummerPotentialGPU,
    "hernquist": HernquistPotentialGPU,
    "isochrone": IsochronePotentialGPU,
    "miyamotonagai": MiyamotoNagaiPotentialGPU,
    "logarithmic": LogHaloPotentialGPU,
    "loghalo": LogHaloPotentialGPU,
    "dehnen": DehnenSphericalPotentialGPU,
    "dehnensph": DehnenSphericalPotentialGPU,
    "diskansatz": DiskAnsatzPotentialGPU,
    "uniformacceleration": UniformAccelerationGPU,
}

# Agama keyword spellings accepted case-insensitively
_KW = {
    "mass": "mass", "scaleradius": "scaleRadius", "scaleheight": "scaleHeight",
    "gamma": "gamma", "velocity": "velocity", "v0": "velocity", "coreradius": "coreRadius",
    "axisratioy": "axisRatioY", "axisratioz": "axisRatioZ",
    "surfacedensity": "surfaceDensity", "innercutoffradius": "innerCutoffRadius",
    "ax": "ax", "ay": "ay", "az": "az", "device": "device",
}


def AnalyticPotentialGPU(type: str, **kwargs) -> PotentialBase:
    """Factory with Agama's constructor syntax, e.g.
    ``AnalyticPotentialGPU(type='NFW', mass=1e12, scaleRadius=20)``.
    Keyword names are matched case-insensitively (the reference lowercases
    them and then fails to map camelCase names back,
    _analytic_potentials.py:1211-1217)."""
    key = type.lower().replace("_", "").replace(" ", "")
    if key not in ANALYTIC_TYPES:
        raise ValueError(f"Unknown analytic potential type '{type}'. "
                         f"Supported: {sorted(ANALYTIC_TYPES)}")
    kw = {}
    for k, v in kwargs.items():
        kk = _KW.get(k.lower().replace("_", ""))
        if kk is None:
            raise TypeError(f"unexpected keyword {k!r} for type {type!r}")
        kw[kk] = v
    return ANALYTIC_TYPES[key](**kw)
