"""Receiver-space objective (reference ``src/tide/workflow/objective.py:45-112``).

``ReceiverObjective(observed, loss).gradient(op.linearize(model))`` is the
model gradient of ``loss(receivers(model), observed)``.  The default loss
is least squares, ``0.5 * ||d - d_obs||^2``, a sum over shots.  For such a
shot-separable loss on a 2-D operator whose primal has not been computed,
the gradient runs chunk by chunk (forward with snapshots, residual,
adjoint) so shot chunks that do not fit device memory together are never
re-run, and the objective value is summed in FP64.  Any other loss goes
through the primal receivers and ``linearized.vjp``.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass

import torch

from ..maxwell.results import EMGradient

ReceiverLoss = Callable[[torch.Tensor, torch.Tensor], torch.Tensor]


def _least_squares(predicted: torch.Tensor, observed: torch.Tensor) -> torch.Tensor:
    return 0.5 * (predicted - observed).square().sum()


_least_squares.shot_separable = True  # type: ignore[attr-defined]


def _separable(loss: ReceiverLoss) -> bool:
    return bool(getattr(loss, "shot_separable", False))


@dataclass(frozen=True)
class ReceiverObjective:
    """A scalar receiver-data objective.  A custom ``loss`` that is a sum of
    per-shot terms may set ``loss.shot_separable = True`` to take the
    chunked path."""

    observed_data: torch.Tensor
    loss: ReceiverLoss = _least_squares

    def _data_gradient(self, predicted: torch.Tensor) -> torch.Tensor:
        data = predicted.detach().requires_grad_(True)
        value = self.loss(data, self.observed_data)
        if value.ndim != 0:
            raise ValueError("Receiver loss must return a scalar tensor.")
        return torch.autograd.grad(value, data)[0]

    def _residual(self, predicted: torch.Tensor, observed: torch.Tensor):
        """(FP64 value, gradient w.r.t. ``predicted``) for one shot chunk."""
        if self.loss is _least_squares:
            diff = predicted - observed
            return 0.5 * (predicted.double() - observed.double()).square().sum(), diff
        data = predicted.detach().requires_grad_(True)
        value = self.loss(data, observed)
        return value.detach().double(), torch.autograd.grad(value, data)[0]

    def value(self, linearized) -> torch.Tensor:
        return self.loss(linearized.primal.receiver_data, self.observed_data)

    def value_and_gradient(self, linearized) -> tuple[torch.Tensor, EMGradient]:
        """The objective value (FP64 on the chunked path) and its gradient."""
        fast = getattr(linearized, "_separable_gradient", None)
        if fast is not None and _separable(self.loss) and linearized._primal is None:
            out = fast(self.observed_data, self._residual)
            if out is not None:
                return out
        grad = linearized.vjp(self._data_gradient(linearized.primal.receiver_data))
        return self.value(linearized).detach(), grad

    def gradient(self, linearized) -> EMGradient:
        return self.value_and_gradient(linearized)[1]

    def hvp(self, linearized, direction, *, mode: str = "full") -> EMGradient:
        if mode not in {"full", "gauss_newton"}:
            raise ValueError("mode must be 'full' or 'gauss_newton'.")
        tangent = linearized.jvp(direction)
        data = linearized.primal.receiver_data.detach().requires_grad_(True)
        value = self.loss(data, self.observed_data)
        g = torch.autograd.grad(value, data, create_graph=True)[0]
        hv = torch.autograd.grad(g, data, tangent.receiver_data)[0]
        gn = linearized.vjp(hv)
        if mode == "gauss_newton":
            return gn
        corr = linearized.second_vjp(direction, g.detach())

        def add(a, b):
            return b if a is None else (a if b is None else a + b)

        return EMGradient(epsilon=add(gn.epsilon, corr.epsilon), sigma=add(gn.sigma, corr.sigma),
                          mu=add(gn.mu, corr.mu))


__all__ = ["ReceiverLoss", "ReceiverObjective"]
