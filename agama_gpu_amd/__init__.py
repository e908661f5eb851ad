"""agama_gpu_amd -- Agama-compatible potential evaluation on AMD MI355X (gfx950).

Drop-in for the reference's ``gpu_potential`` module: the ``PotentialGPU``
factory and potential classes with Agama's API
(potential / force / density / forceDeriv / evalDeriv / eval), FP64, on
PyTorch-ROCm tensors, evaluated by hand-written gfx950 HIP kernels through
the C ABI in ``include/agama_gpu.h``.
"""
from .coefs import CylSplineCoefs, MultipoleCoefs, read_coefs, write_coefs
from .multipole_prep import G_AGAMA, INV_4PIG
from .base import PotentialBase
from .multipole import MultipolePotentialGPU
from .cylspline import CylSplinePotentialGPU
from .composite import (CompositePotentialGPU, EvolvingPotentialGPU, ScaledPotentialGPU,
                        ShiftedPotentialGPU)
from .analytic import (AnalyticPotentialGPU, DehnenSphericalPotentialGPU, DiskAnsatzPotentialGPU,
                       HernquistPotentialGPU, IsochronePotentialGPU, LogHaloPotentialGPU,
                       MiyamotoNagaiPotentialGPU, NFWPotentialGPU, PlummerPotentialGPU,
                       UniformAccelerationGPU)
from .factory import PotentialGPU
from .batch import MultipoleEnsembleGPU
from .integrate import LeapfrogGPU
from .autograd import FittableCylSpline, FittableMultipole, FittableShiftedMultipole, SplineCenter

__all__ = [
    "PotentialGPU", "PotentialBase", "MultipolePotentialGPU", "CylSplinePotentialGPU",
    "CompositePotentialGPU",
    "EvolvingPotentialGPU", "ShiftedPotentialGPU", "ScaledPotentialGPU",
    "AnalyticPotentialGPU", "NFWPotentialGPU", "PlummerPotentialGPU", "HernquistPotentialGPU",
    "DehnenSphericalPotentialGPU", "IsochronePotentialGPU", "MiyamotoNagaiPotentialGPU",
    "LogHaloPotentialGPU", "DiskAnsatzPotentialGPU", "UniformAccelerationGPU",
    "MultipoleCoefs", "CylSplineCoefs", "read_coefs", "write_coefs", "G_AGAMA", "INV_4PIG",
    "MultipoleEnsembleGPU", "LeapfrogGPU", "FittableMultipole", "FittableShiftedMultipole",
    "SplineCenter", "FittableCylSpline",
]
__version__ = "0.1.0"
