"""TIDE's Maxwell FDTD propagators, MI355X-native (HIP on gfx950).

Public surface for the propagation hot path, mirroring the reference
package: ``MaxwellTM`` / ``Maxwell3D`` operators over ``EMModel`` with
``Experiment`` / ``Acquisition`` / ``Discretization`` / ``CPML`` inputs,
``StorageOptions`` snapshot policy (device, host-staged or disk),
wavefield callbacks, ``linearize().vjp``, the ``workflow``
shot-batching and receiver-objective helpers; plus ``ricker`` and
``cfl_condition``.  The compute path is the HIP library in ``_lib``; there
is no CPU fallback.
"""

from . import callbacks, cfl, padding, staggered, storage, utils, wavelets
from .callbacks import Callback, CallbackState
from .cfl import cfl_condition
from .core import BackendPreference, FallbackPolicy, StorageMode, StorageOptions
from .dispersion import DebyeDispersion
from .maxwell import (
    CPML,
    Acquisition,
    Discretization,
    EM3DState,
    EMDirection,
    EMGradient,
    EMModel,
    ExecutionOptions,
    Experiment,
    ForwardResult,
    Maxwell3D,
    MaxwellTM,
    Observers,
    SourceConvention,
    TangentResult,
    TMState,
)
from .utils import C0, EP0, MU0
from .wavelets import gaussian, gaussian_derivative, ricker
from . import workflow  # noqa: E402  (imports the operators above)

__all__ = [
    "DebyeDispersion",
    "Acquisition",
    "BackendPreference",
    "C0",
    "CPML",
    "Callback",
    "CallbackState",
    "Discretization",
    "EM3DState",
    "EMDirection",
    "EMGradient",
    "EMModel",
    "EP0",
    "ExecutionOptions",
    "Experiment",
    "FallbackPolicy",
    "ForwardResult",
    "MU0",
    "Maxwell3D",
    "MaxwellTM",
    "Observers",
    "SourceConvention",
    "StorageMode",
    "StorageOptions",
    "TMState",
    "TangentResult",
    "callbacks",
    "cfl",
    "cfl_condition",
    "gaussian",
    "gaussian_derivative",
    "padding",
    "ricker",
    "staggered",
    "storage",
    "utils",
    "wavelets",
    "workflow",
]
