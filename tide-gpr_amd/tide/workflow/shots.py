"""Shot mini-batch helpers (reference ``src/tide/workflow/shots.py:1-232``).

Shots are independent: a batch is a contiguous range of shot indices, its
source and receiver tensors are the matching rows of the full acquisition,
and receivers of several batches concatenate along the shot axis.  The
operator itself already splits shots into chunks that fit device memory;
these helpers serve callers that batch shots in their own loops (for
example stochastic shot selection).
"""

from __future__ import annotations

from collections.abc import Callable, Iterable, Sequence
from dataclasses import dataclass
from typing import Any

import torch

from ..maxwell.contracts import Acquisition, EMModel, Experiment

ModelOutput = torch.Tensor | Sequence[torch.Tensor]
ReceiverSelector = Callable[[ModelOutput], torch.Tensor]


@dataclass(frozen=True)
class ShotBatch:
    """Source and receiver tensors of one shot mini-batch."""

    source_amplitude: torch.Tensor | None
    source_location: torch.Tensor | None
    receiver_location: torch.Tensor | None


def split_shots(n_shots: int, batch_size: int,
                device: torch.device | str | None = None) -> list[torch.Tensor]:
    """Contiguous batches ``[0, b), [b, 2b), ...`` of shot indices."""
    n_shots, batch_size = int(n_shots), int(batch_size)
    if n_shots < 0:
        raise ValueError("n_shots must be non-negative.")
    if batch_size <= 0:
        raise ValueError("batch_size must be positive.")
    return list(torch.arange(n_shots, device=device, dtype=torch.long).split(batch_size))


def _axis(ndim: int, dim: int) -> int:
    if not -ndim <= int(dim) < ndim:
        raise ValueError(f"shot_dim={dim} is out of bounds for tensor with ndim={ndim}.")
    return int(dim) % ndim


def index_shots(tensor: torch.Tensor | None, shot_indices: torch.Tensor, *,
                shot_dim: int = 0) -> torch.Tensor | None:
    """The rows ``shot_indices`` of ``tensor`` along ``shot_dim``."""
    if tensor is None:
        return None
    if shot_indices.ndim != 1:
        raise ValueError("shot_indices must be a 1D tensor.")
    idx = shot_indices.to(device=tensor.device, dtype=torch.long)
    return torch.index_select(tensor, _axis(tensor.ndim, shot_dim), idx)


def take_shot_batch(*, source_amplitude: torch.Tensor | None,
                    source_location: torch.Tensor | None,
                    receiver_location: torch.Tensor | None, shot_indices: torch.Tensor,
                    shot_dim: int = 0) -> ShotBatch:
    return ShotBatch(*(index_shots(t, shot_indices, shot_dim=shot_dim)
                       for t in (source_amplitude, source_location, receiver_location)))


def infer_receiver_shot_dim(receiver: torch.Tensor) -> int:
    """Shot axis of ``[nt, S, R]`` or ``[nt, B, S, R]`` receiver data."""
    if receiver.ndim not in (3, 4):
        raise ValueError(
            "receiver chunks must be shaped [nt, S, R] or [nt, B, S, R] "
            f"when receiver_shot_dim is not provided, got {tuple(receiver.shape)}."
        )
    return receiver.ndim - 2


def merge_receiver_batches(chunks: Iterable[torch.Tensor], *,
                           shot_dim: int | None = None) -> torch.Tensor:
    parts = list(chunks)
    if not parts:
        raise ValueError("chunks must contain at least one receiver tensor.")
    dim = infer_receiver_shot_dim(parts[0]) if shot_dim is None else shot_dim
    dim = _axis(parts[0].ndim, dim)
    return parts[0] if len(parts) == 1 else torch.cat(parts, dim=dim)


def _run_kernel_shot_batches(solver: Callable[..., ModelOutput], *, n_shots: int,
                             batch_size: int, source_amplitude: torch.Tensor | None = None,
                             source_location: torch.Tensor | None = None,
                             receiver_location: torch.Tensor | None = None, shot_dim: int = 0,
                             receiver_shot_dim: int | None = None,
                             receiver_selector: ReceiverSelector | None = None,
                             device: torch.device | str | None = None,
                             **model_kwargs: Any) -> torch.Tensor:
    """Run a functional solver (``maxwelltm``-like) batch by batch and
    concatenate its receivers (the last output by default)."""
    pick = receiver_selector or (lambda out: out if isinstance(out, torch.Tensor) else out[-1])
    tensors = (source_amplitude, source_location, receiver_location)
    dev = device if device is not None else next((t.device for t in tensors if t is not None), None)
    chunks = []
    for idx in split_shots(n_shots, batch_size, dev):
        b = take_shot_batch(source_amplitude=source_amplitude, source_location=source_location,
                            receiver_location=receiver_location, shot_indices=idx,
                            shot_dim=shot_dim)
        chunks.append(pick(solver(source_amplitude=b.source_amplitude,
                                  source_location=b.source_location,
                                  receiver_location=b.receiver_location, **model_kwargs)))
    return merge_receiver_batches(chunks, shot_dim=receiver_shot_dim)


def batch_operator(operator, shot_indices: torch.Tensor):
    """A copy of ``operator`` restricted to ``shot_indices``."""
    exp = operator.experiment
    acq = exp.acquisition
    b = take_shot_batch(source_amplitude=exp.source_amplitude,
                        source_location=acq.source_location,
                        receiver_location=acq.receiver_location, shot_indices=shot_indices)
    sub = Experiment(
        acquisition=Acquisition(source_location=b.source_location,
                                receiver_location=b.receiver_location),
        source_amplitude=b.source_amplitude, nt=exp.nt,
        source_component=exp.source_component, receiver_component=exp.receiver_component,
        source_convention=exp.source_convention,
        frequency_taper_fraction=exp.frequency_taper_fraction,
        time_padding_fraction=exp.time_padding_fraction, time_taper=exp.time_taper,
    )
    return type(operator)(operator.discretization, sub, execution=operator.execution,
                          storage=operator.storage,
                          model_gradient_sampling_interval=operator.model_gradient_sampling_interval)


def run_shot_batches(operator, model: EMModel, *, batch_size: int,
                     receiver_shot_dim: int | None = None) -> torch.Tensor:
    """Receivers of ``operator(model)`` computed ``batch_size`` shots at a time."""
    chunks = [batch_operator(operator, idx)(model).receiver_data
              for idx in split_shots(operator.experiment.acquisition.n_shots, batch_size,
                                     model.epsilon.device)]
    return merge_receiver_batches(chunks, shot_dim=receiver_shot_dim)


__all__ = [
    "ShotBatch",
    "batch_operator",
    "infer_receiver_shot_dim",
    "index_shots",
    "merge_receiver_batches",
    "run_shot_batches",
    "split_shots",
    "take_shot_batch",
]
