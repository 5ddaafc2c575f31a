"""Public value objects of the Maxwell operators.

Field names, defaults and validation follow the reference's public contract
(``src/tide/maxwell/contracts.py:16-201``) so user code constructs the same
objects.  Python 3.10 has no ``enum.StrEnum``; ``SourceConvention`` is a
``str`` enum with the same values.
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Any, Literal

import torch

from ..core import BackendPreference, FallbackPolicy


class SourceConvention(str, Enum):
    """How the supplied source amplitude enters the discrete update."""

    PHYSICAL_CURRENT = "physical_current"
    FIELD_INCREMENT = "field_increment"


@dataclass(frozen=True)
class CPML:
    """Convolutional PML width in cells (one value, or per side)."""

    width: int | tuple[int, ...] = 20

    def __post_init__(self) -> None:
        widths = (self.width,) if isinstance(self.width, int) else self.width
        if not widths or any(w < 0 for w in widths):
            raise ValueError("CPML widths must be non-negative.")


@dataclass(frozen=True)
class Discretization:
    """Grid spacing (m), time step (s), stencil order and boundary."""

    spacing: float | tuple[float, ...]
    dt: float
    stencil: int = 2
    boundary: CPML = CPML()
    max_velocity: float | None = None

    def __post_init__(self) -> None:
        spacing = (self.spacing,) if isinstance(self.spacing, (int, float)) else self.spacing
        if not spacing or any(v <= 0 for v in spacing):
            raise ValueError("Grid spacing must be positive.")
        if self.dt <= 0:
            raise ValueError("dt must be positive.")
        if self.stencil not in {2, 4, 6, 8}:
            raise ValueError("stencil must be one of 2, 4, 6, or 8.")
        if self.max_velocity is not None and self.max_velocity <= 0:
            raise ValueError("max_velocity must be positive when provided.")


@dataclass(frozen=True)
class Acquisition:
    """Source and receiver cell coordinates, ``[shots, n, ndim]``."""

    source_location: torch.Tensor | None
    receiver_location: torch.Tensor | None

    @property
    def n_shots(self) -> int:
        for t in (self.source_location, self.receiver_location):
            if t is not None:
                return int(t.shape[0])
        return 1

    @property
    def n_receivers(self) -> int:
        return 0 if self.receiver_location is None else int(self.receiver_location.shape[1])

    @property
    def spatial_ndim(self) -> int | None:
        for t in (self.source_location, self.receiver_location):
            if t is not None:
                return int(t.shape[-1])
        return None


@dataclass(frozen=True)
class Experiment:
    """Acquisition, source wavelets ``[shots, n_src, nt]`` and time axis."""

    acquisition: Acquisition
    source_amplitude: torch.Tensor | None
    nt: int | None = None
    source_component: str = "ey"
    receiver_component: str = "ey"
    source_convention: SourceConvention = SourceConvention.PHYSICAL_CURRENT
    frequency_taper_fraction: float = 0.0
    time_padding_fraction: float = 0.0
    time_taper: bool = False

    def __post_init__(self) -> None:
        if self.nt is not None and self.nt <= 0:
            raise ValueError("nt must be positive when provided.")
        if self.source_amplitude is None and self.nt is None:
            raise ValueError("nt is required when source_amplitude is absent.")
        if not self.source_component or not self.receiver_component:
            raise ValueError("Source and receiver components must be non-empty.")
        for name, v in (
            ("frequency_taper_fraction", self.frequency_taper_fraction),
            ("time_padding_fraction", self.time_padding_fraction),
        ):
            if not 0.0 <= v <= 1.0:
                raise ValueError(f"{name} must be between zero and one.")


@dataclass(frozen=True)
class EMModel:
    """Relative permittivity, conductivity (S/m) and relative permeability."""

    epsilon: torch.Tensor
    sigma: torch.Tensor
    mu: torch.Tensor
    dispersion: Any | None = None

    def __post_init__(self) -> None:
        if self.sigma.shape != self.epsilon.shape or self.mu.shape != self.epsilon.shape:
            raise ValueError("epsilon, sigma, and mu must have the same shape.")
        if self.sigma.device != self.epsilon.device or self.mu.device != self.epsilon.device:
            raise ValueError("epsilon, sigma, and mu must be on the same device.")
        if self.sigma.dtype != self.epsilon.dtype or self.mu.dtype != self.epsilon.dtype:
            raise ValueError("epsilon, sigma, and mu must have the same dtype.")


@dataclass(frozen=True)
class EMDirection:
    """A direction in model space."""

    epsilon: torch.Tensor | None = None
    sigma: torch.Tensor | None = None
    mu: torch.Tensor | None = None

    def validate_for(self, model: "EMModel") -> None:
        if self.epsilon is None and self.sigma is None and self.mu is None:
            raise ValueError("At least one model direction must be provided.")
        for name in ("epsilon", "sigma", "mu"):
            d, m = getattr(self, name), getattr(model, name)
            if d is not None and (d.shape != m.shape or d.device != m.device or d.dtype != m.dtype):
                raise ValueError(f"The {name} direction must match the corresponding model tensor.")


@dataclass(frozen=True)
class Observers:
    """Wavefield observers (callbacks) for forward/reverse propagation."""

    forward: Any | None = None
    backward: Any | None = None
    frequency: int = 1

    def __post_init__(self) -> None:
        if self.frequency <= 0:
            raise ValueError("Observer frequency must be positive.")


@dataclass(frozen=True)
class ExecutionOptions:
    """Backend policy.  This build has a single backend (HIP on MI355X);
    ``backend='reference'`` is rejected at call time."""

    backend: BackendPreference = BackendPreference.AUTO
    fallback: FallbackPolicy = FallbackPolicy.ERROR
    reference_mode: Literal["eager", "jit", "compile"] = "eager"
    n_threads: int | None = None

    def __post_init__(self) -> None:
        if self.n_threads is not None and self.n_threads < 0:
            raise ValueError("n_threads must be non-negative when provided.")

    @property
    def legacy_backend_request(self) -> bool | str:
        if self.backend is BackendPreference.NATIVE:
            return "native"
        if self.backend is BackendPreference.REFERENCE:
            return self.reference_mode
        return False


__all__ = [
    "Acquisition",
    "CPML",
    "Discretization",
    "EMDirection",
    "EMModel",
    "ExecutionOptions",
    "Experiment",
    "Observers",
    "SourceConvention",
]
