"""CPML profile assembly for the staggered grid.

``set_pml_profiles`` returns the 12 per-axis profiles the 2-D kernels read:
``[ay, ayh, ax, axh, by, byh, bx, bxh, ky, kyh, kx, kxh]`` shaped for
broadcasting over ``[shot, y, x]`` (reference ``src/tide/staggered.py:6-134``).
The inner strip edges are at ``fd_pad + pml_width`` from the low side and
``n - 1 - fd_pad - pml_width`` from the high side; the deepest strip of any
side sets the common ``max_pml`` length.
"""

from __future__ import annotations

import torch

from . import utils


def _pml_edges(n: int, width_lo: int, width_hi: int, pad_lo: int, pad_hi: int) -> list[float]:
    return [pad_lo + width_lo, n - 1 - pad_hi - width_hi]


def set_pml_profiles(
    pml_width: list[int],
    accuracy: int,
    fd_pad: list[int],
    dt: float,
    grid_spacing: list[float],
    max_vel: float,
    dtype: torch.dtype,
    device: torch.device,
    pml_freq: float,
    ny: int,
    nx: int,
    eps_scale: float = utils.EP0,
) -> list[torch.Tensor]:
    del accuracy
    edges_y = _pml_edges(ny, pml_width[0], pml_width[1], fd_pad[0], fd_pad[1])
    edges_x = _pml_edges(nx, pml_width[2], pml_width[3], fd_pad[2], fd_pad[3])
    max_pml = max(
        pml_width[0] * grid_spacing[0],
        pml_width[1] * grid_spacing[0],
        pml_width[2] * grid_spacing[1],
        pml_width[3] * grid_spacing[1],
    )

    def axis(widths, edges, n, spacing, half):
        setup = utils.setup_pml_half if half else utils.setup_pml
        return setup(
            widths,
            edges,
            max_pml,
            dt,
            n,
            max_vel,
            dtype,
            device,
            pml_freq,
            start=0.0,
            grid_spacing=spacing,
            eps_scale=eps_scale,
        )

    ay, by, ky = axis(pml_width[:2], edges_y, ny, grid_spacing[0], False)
    ax, bx, kx = axis(pml_width[2:], edges_x, nx, grid_spacing[1], False)
    ayh, byh, kyh = axis(pml_width[:2], edges_y, ny, grid_spacing[0], True)
    axh, bxh, kxh = axis(pml_width[2:], edges_x, nx, grid_spacing[1], True)

    def col(v):  # varies along y
        return v[None, :, None]

    def row(v):  # varies along x
        return v[None, None, :]

    return [
        col(ay), col(ayh), row(ax), row(axh),
        col(by), col(byh), row(bx), row(bxh),
        col(ky), col(kyh), row(kx), row(kxh),
    ]


def set_pml_profiles_3d(
    pml_width: list[int],
    fd_pad: list[int],
    dt: float,
    grid_spacing: list[float],
    dtype: torch.dtype,
    device: torch.device,
    nz: int,
    ny: int,
    nx: int,
    eps_scale: float = utils.EP0,
) -> dict[str, torch.Tensor]:
    """Per-axis 1-D CPML profiles for a 3-D grid ``[nz, ny, nx]``.

    Keys ``a{ax}``, ``b{ax}``, ``k{ax}`` and their half-point ``…h`` variants
    for ``ax`` in ``z, y, x`` (reference ``src/tide/staggered.py:188-354``).
    """
    sizes = {"z": nz, "y": ny, "x": nx}
    max_pml = max(
        pml_width[0] * grid_spacing[0],
        pml_width[1] * grid_spacing[0],
        pml_width[2] * grid_spacing[1],
        pml_width[3] * grid_spacing[1],
        pml_width[4] * grid_spacing[2],
        pml_width[5] * grid_spacing[2],
    )
    out: dict[str, torch.Tensor] = {}
    for d, name in enumerate(("z", "y", "x")):
        n = sizes[name]
        widths = pml_width[2 * d : 2 * d + 2]
        edges = _pml_edges(n, widths[0], widths[1], fd_pad[2 * d], fd_pad[2 * d + 1])
        for half, setup in ((False, utils.setup_pml), (True, utils.setup_pml_half)):
            a, b, k = setup(
                widths,
                edges,
                max_pml,
                dt,
                n,
                0.0,
                dtype,
                device,
                0.5 / dt,
                start=0.0,
                grid_spacing=grid_spacing[d],
                eps_scale=eps_scale,
            )
            suffix = "h" if half else ""
            out[f"a{name}{suffix}"] = a
            out[f"b{name}{suffix}"] = b
            out[f"k{name}{suffix}"] = k
    return out


__all__ = ["set_pml_profiles", "set_pml_profiles_3d"]
