"""CFL time-step helper (reference ``src/tide/cfl.py:6-63``)."""

from __future__ import annotations

import math
import warnings


def cfl_condition(
    grid_spacing: float | list[float],
    dt: float,
    max_vel: float,
    c_max: float = 1.0,
    eps: float = 1e-15,
) -> tuple[float, int]:
    """Return ``(inner_dt, step_ratio)``: the largest stable
    ``dt_max = c_max / sqrt(sum 1/h^2) / v_max`` and the number of internal
    steps per user step, ``ceil(|dt| / dt_max)``."""
    if isinstance(grid_spacing, (int, float)):
        spacing = [float(grid_spacing), float(grid_spacing)]
    else:
        spacing = list(grid_spacing)
    if max_vel <= 0:
        raise ValueError("max_vel must be positive")
    max_dt = (c_max / math.sqrt(sum(1 / h**2 for h in spacing)) / (max_vel**2 + eps)) * max_vel
    step_ratio = math.ceil(abs(dt) / max_dt)
    inner_dt = dt / step_ratio
    if step_ratio >= 2:
        warnings.warn(
            f"CFL condition requires {step_ratio} internal time steps per "
            f"user time step (dt={dt}, inner_dt={inner_dt}). Consider using "
            "a smaller dt or coarser grid.",
            stacklevel=2,
        )
    return inner_dt, step_ratio


__all__ = ["cfl_condition"]
