"""Debye dispersion parameters (reference ``src/tide/dispersion.py``)."""

from __future__ import annotations

from dataclasses import dataclass

import torch


@dataclass(frozen=True)
class DebyeDispersion:
    """Debye poles: ``delta_epsilon`` (relative permittivity step) and
    ``tau`` (relaxation time, s), each a scalar, a model-shaped tensor or a
    ``[n_poles, *model_shape]`` tensor; scalars and model-shaped tensors
    are one pole."""

    delta_epsilon: torch.Tensor | float
    tau: torch.Tensor | float


__all__ = ["DebyeDispersion"]
