"""Model and wavefield padding (reference ``src/tide/padding.py``).

Models are replicate-padded into the CPML and finite-difference halo, so the
gradient of a padded cell folds back onto the nearest edge cell through
``F.pad``'s backward; wavefields are zero-padded.  Pads are given per side as
``[y0, y1, x0, x1]`` (2-D) or ``[z0, z1, y0, y1, x0, x1]`` (3-D).
"""

from __future__ import annotations

import torch


def reverse_pad(pad: list[int]) -> list[int]:
    """``[y0, y1, x0, x1]`` -> ``[x0, x1, y0, y1]`` (last dimension first)."""
    pairs = [pad[i : i + 2] for i in range(0, len(pad), 2)]
    return [v for pair in reversed(pairs) for v in pair]


def create_or_pad(
    tensor: torch.Tensor,
    pad: int | list[int],
    device: torch.device,
    dtype: torch.dtype,
    size: tuple[int, ...],
    mode: str = "constant",
) -> torch.Tensor:
    """Zeros of ``size`` for an empty tensor, else ``F.pad(tensor, pad, mode)``."""
    if isinstance(pad, int):
        spatial = len(size) - 1 if len(size) > 2 else len(size)
        pad = [pad] * (2 * spatial)
    if tensor.numel() == 0:
        return torch.zeros(size, device=device, dtype=dtype)
    if max(pad) == 0:
        return tensor.clone()
    spatial = len(pad) // 2
    # Non-constant modes need a leading batch/channel dimension.
    squeeze = mode != "constant" and tensor.ndim in (spatial, spatial + 1)
    src = tensor.unsqueeze(0) if squeeze else tensor
    out = torch.nn.functional.pad(src, reverse_pad(pad), mode=mode)
    return out.squeeze(0) if squeeze else out


def zero_interior(
    tensor: torch.Tensor,
    fd_pad: int | list[int],
    pml_width: list[int],
    dim: int,
) -> torch.Tensor:
    """Zero the non-CPML band of a CPML memory along spatial axis ``dim``."""
    shape = tensor.shape[1:]
    ndim = len(shape)
    if ndim not in (2, 3):
        raise ValueError(f"zero_interior expects 2D or 3D spatial tensors, got ndim={ndim}")
    if isinstance(fd_pad, int):
        fd_pad = [fd_pad] * (2 * ndim)
    if len(fd_pad) != 2 * ndim or len(pml_width) != 2 * ndim:
        raise ValueError("fd_pad and pml_width need two entries per spatial axis")
    if not 0 <= dim < ndim:
        raise ValueError(f"dim must be in [0, {ndim - 1}], got {dim}.")
    lo = fd_pad[2 * dim] + pml_width[2 * dim]
    hi = shape[dim] - pml_width[2 * dim + 1] - fd_pad[2 * dim + 1]
    index = [slice(None)] * tensor.ndim
    index[dim + 1] = slice(lo, hi)
    tensor[tuple(index)].fill_(0)
    return tensor


__all__ = ["create_or_pad", "reverse_pad", "zero_interior"]
