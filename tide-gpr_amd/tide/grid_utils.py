"""Normalisation of grid spacing and CPML widths (reference
``src/tide/grid_utils.py:9-95``)."""

from __future__ import annotations

from collections.abc import Sequence


def _normalize_grid_spacing_2d(grid_spacing: float | Sequence[float]) -> list[float]:
    """-> ``[dy, dx]``."""
    if isinstance(grid_spacing, (int, float)):
        return [float(grid_spacing)] * 2
    vals = list(grid_spacing)
    if len(vals) == 1:
        return [float(vals[0])] * 2
    if len(vals) != 2:
        raise ValueError(f"2D grid_spacing must have length 1 or 2, got {len(vals)}.")
    return [float(vals[0]), float(vals[1])]


def _normalize_grid_spacing_3d(grid_spacing: float | Sequence[float]) -> list[float]:
    """-> ``[dz, dy, dx]``."""
    if isinstance(grid_spacing, (int, float)):
        return [float(grid_spacing)] * 3
    vals = list(grid_spacing)
    if len(vals) == 1:
        return [float(vals[0])] * 3
    if len(vals) != 3:
        raise ValueError(f"3D grid_spacing must have length 1 or 3, got {len(vals)}.")
    return [float(v) for v in vals]


def _normalize_pml_width_2d(pml_width: int | Sequence[int]) -> list[int]:
    """-> ``[top, bottom, left, right]``."""
    if isinstance(pml_width, int):
        return [pml_width] * 4
    vals = list(pml_width)
    if len(vals) == 1:
        return [int(vals[0])] * 4
    if len(vals) == 2:
        return [int(vals[0]), int(vals[0]), int(vals[1]), int(vals[1])]
    if len(vals) != 4:
        raise ValueError(f"2D pml_width must have length 1, 2, or 4, got {len(vals)}.")
    return [int(v) for v in vals]


def _normalize_pml_width_3d(pml_width: int | Sequence[int]) -> list[int]:
    """-> ``[z0, z1, y0, y1, x0, x1]``."""
    if isinstance(pml_width, int):
        return [pml_width] * 6
    vals = list(pml_width)
    if len(vals) == 1:
        return [int(vals[0])] * 6
    if len(vals) == 3:
        return [int(vals[0]), int(vals[0]), int(vals[1]), int(vals[1]), int(vals[2]), int(vals[2])]
    if len(vals) != 6:
        raise ValueError(f"3D pml_width must have length 1, 3, or 6, got {len(vals)}.")
    return [int(v) for v in vals]
