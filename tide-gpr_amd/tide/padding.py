"""Model and wavefield padding (reference ``src/tide/padding.py``).

Models are replicate-padded into the CPML and finite-difference halo, so the
gradient of a padded cell folds back onto the nearest edge cell; wavefields
are zero-padded.  Pads are given per side as ``[y0, y1, x0, x1]`` (2-D) or
``[z0, z1, y0, y1, x0, x1]`` (3-D).

The fold is done here with slices and reductions rather than by
``F.pad``'s backward, which adds with atomics on the GPU and so is not the
same from run to run; with this fold the operators' gradients are bitwise
repeatable.
"""

from __future__ import annotations

import torch


def fold_replicate(g_p: torch.Tensor, pad: list[int], spatial: int) -> torch.Tensor:
    """Adjoint of replicate padding of the last ``spatial`` dimensions by
    ``pad`` (``[lo, hi]`` per dimension, first dimension first): each padded
    cell's gradient is added to the edge cell it copies, the slab of each
    side summed by one reduction (fixed order, no atomics)."""
    out = g_p
    for d in range(spatial):
        lo, hi = int(pad[2 * d]), int(pad[2 * d + 1])
        if lo == 0 and hi == 0:
            continue
        axis = out.ndim - spatial + d
        n = out.shape[axis] - lo - hi
        center = out.narrow(axis, lo, n).clone()
        if lo:
            center.narrow(axis, 0, 1).add_(out.narrow(axis, 0, lo).sum(axis, keepdim=True))
        if hi:
            center.narrow(axis, n - 1, 1).add_(out.narrow(axis, lo + n, hi).sum(axis, keepdim=True))
        out = center
    return out


class _ReplicatePad(torch.autograd.Function):
    """Replicate padding with the deterministic fold as its backward."""

    @staticmethod
    def forward(ctx, tensor, pad_rev, pad, spatial):
        ctx.pad = pad
        ctx.spatial = spatial
        return torch.nn.functional.pad(tensor, pad_rev, mode="replicate")

    @staticmethod
    def backward(ctx, grad):
        return fold_replicate(grad, ctx.pad, ctx.spatial), None, None, None


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
    if mode == "replicate":
        out = _ReplicatePad.apply(src, reverse_pad(pad), list(pad), spatial)
    else:
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


__all__ = ["create_or_pad", "fold_replicate", "reverse_pad", "zero_interior"]
