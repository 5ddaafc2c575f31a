from .contracts import (
    CPML,
    Acquisition,
    Discretization,
    EMDirection,
    EMModel,
    ExecutionOptions,
    Experiment,
    Observers,
    SourceConvention,
)
from .operators import Maxwell3D, MaxwellTM
from .results import EM3DState, EMGradient, ForwardResult, TangentResult, TMState
from .tm2d_born import borntm

__all__ = [
    "Acquisition",
    "CPML",
    "Discretization",
    "EM3DState",
    "EMDirection",
    "EMGradient",
    "EMModel",
    "ExecutionOptions",
    "Experiment",
    "ForwardResult",
    "Maxwell3D",
    "MaxwellTM",
    "Observers",
    "SourceConvention",
    "TMState",
    "TangentResult",
    "borntm",
]
