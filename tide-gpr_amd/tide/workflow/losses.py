"""Receiver-loss helpers for shot-batched loops (reference
``src/tide/workflow/losses.py:21-66, 418-463``)."""

from __future__ import annotations

from collections.abc import Callable, Iterable
from typing import Literal

import torch
import torch.nn.functional as F

from .shots import index_shots, infer_receiver_shot_dim

LossNormalization = Literal["batch", "all", "sum"]


def take_receiver_batch(receiver: torch.Tensor, shot_indices: torch.Tensor, *,
                        shot_dim: int | None = None) -> torch.Tensor:
    dim = infer_receiver_shot_dim(receiver) if shot_dim is None else shot_dim
    return index_shots(receiver, shot_indices, shot_dim=dim)


def receiver_mse_loss(predicted: torch.Tensor, observed: torch.Tensor,
                      shot_indices: torch.Tensor, *, shot_dim: int | None = None,
                      normalization: LossNormalization = "batch") -> torch.Tensor:
    """Mean squared error against the observed shots ``shot_indices``:
    ``batch`` averages over the batch, ``all`` divides the squared sum by
    every observed sample, ``sum`` does not normalise."""
    target = take_receiver_batch(observed, shot_indices, shot_dim=shot_dim)
    if normalization == "batch":
        return F.mse_loss(predicted, target)
    sq = (predicted - target).square().sum()
    if normalization == "all":
        return sq / observed.numel()
    if normalization == "sum":
        return sq
    raise ValueError("normalization must be 'batch', 'all', or 'sum'.")


def backward_shot_batches(loss_fn: Callable[[torch.Tensor], torch.Tensor],
                          shot_batches: Iterable[torch.Tensor], *,
                          zero_grad: Callable[[], None] | None = None,
                          zero_each_batch: bool = False,
                          after_backward: Callable[[torch.Tensor, torch.Tensor], None] | None = None,
                          retain_graph: bool = False, create_graph: bool = False) -> float:
    """``loss_fn(batch).backward()`` for every batch; returns the summed
    loss (kept on the device until the last batch is launched)."""
    if zero_grad is not None and not zero_each_batch:
        zero_grad()
    total: torch.Tensor | None = None
    for idx in shot_batches:
        if zero_grad is not None and zero_each_batch:
            zero_grad()
        loss = loss_fn(idx)
        if not isinstance(loss, torch.Tensor):
            raise TypeError("loss_fn must return a torch.Tensor.")
        if loss.ndim != 0:
            raise ValueError("loss_fn must return a scalar loss tensor.")
        loss.backward(retain_graph=retain_graph, create_graph=create_graph)
        if after_backward is not None:
            after_backward(idx, loss)
        total = loss.detach() if total is None else total + loss.detach()
    if total is None:
        raise ValueError("shot_batches must contain at least one batch.")
    return float(total)


__all__ = ["LossNormalization", "backward_shot_batches", "receiver_mse_loss",
           "take_receiver_batch"]
