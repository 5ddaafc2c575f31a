"""Jacobian products of the Maxwell operators (reference
``src/tide/maxwell/linearization.py:36-229``).

``op.linearize(model)`` runs the primal once with a differentiable model and
``.vjp(cotangent)`` applies ``J(model)^T`` to a receiver cotangent through
the hand-written adjoint.  For the 2-D operator ``.jvp(direction)`` applies
``J(model)`` by Born (tangent-linear) propagation (``tm2d_born.borntm``) and
``.second_vjp(direction, cotangent)`` the model gradient of
``<J(model) direction, cotangent>`` through the native adjoint of the Born
recurrence.  The 3-D tangent is not built.
"""

from __future__ import annotations

from collections.abc import Iterable

import torch

from ..core import StorageOptions
from .contracts import EMDirection, EMModel, SourceConvention
from .results import EM3DState, EMGradient, ForwardResult, TangentResult, TMState

_TARGETS = ("epsilon", "sigma", "mu")


class _LinearizedMaxwell:
    def __init__(self, operator, model: EMModel, *, storage: StorageOptions,
                 targets: Iterable[str] = ("epsilon", "sigma")) -> None:
        target_set = frozenset(targets)
        unknown = target_set - set(_TARGETS)
        if unknown:
            raise ValueError(f"Unknown linearization targets: {sorted(unknown)!r}.")
        if not target_set:
            raise ValueError("At least one linearization target is required.")
        if "mu" in target_set:
            raise NotImplementedError("gradients with respect to mu are not part of this build")
        if storage != operator.storage:
            operator = type(operator)(
                operator.discretization, operator.experiment,
                execution=operator.execution, storage=storage,
                model_gradient_sampling_interval=operator.model_gradient_sampling_interval,
            )
        self.operator = operator
        self.model = EMModel(
            epsilon=model.epsilon.detach().requires_grad_("epsilon" in target_set),
            sigma=model.sigma.detach().requires_grad_("sigma" in target_set),
            mu=model.mu.detach(),
            dispersion=model.dispersion,
        )
        self.storage = storage
        self.targets = target_set
        self._primal: ForwardResult | None = None
        self._closed = False

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def close(self) -> None:
        self._primal = None
        self._closed = True

    @property
    def primal(self) -> ForwardResult:
        if self._closed:
            raise RuntimeError("The linearized Maxwell operator is closed.")
        if self._primal is None:
            self._primal = self.operator.forward(self.model)
        return self._primal

    def _parameters(self) -> tuple[torch.Tensor, ...]:
        return tuple(getattr(self.model, n) for n in _TARGETS if n in self.targets)

    def vjp(self, cotangent: torch.Tensor) -> EMGradient:
        """``J(model)^T cotangent`` as an :class:`EMGradient`."""
        receiver_data = self.primal.receiver_data
        if cotangent.shape != receiver_data.shape:
            raise ValueError("Receiver cotangent must match primal receiver data shape.")
        grads = torch.autograd.grad(
            receiver_data, self._parameters(), cotangent, retain_graph=True, allow_unused=True
        )
        values = {n: None for n in _TARGETS}
        for n, g in zip((n for n in _TARGETS if n in self.targets), grads):
            values[n] = g
        return EMGradient(**values)

    def jvp(self, direction):
        raise NotImplementedError("tangent products of this operator are not part of this build")

    def second_vjp(self, *args, **kwargs):
        raise NotImplementedError("second-order products are not part of this build")


class LinearizedMaxwellTM(_LinearizedMaxwell):
    def jvp(self, direction: EMDirection) -> TangentResult:
        """``J(model) direction``: receivers and final state of the Born
        scattered field (reference ``linearization.py:178-229``).  The
        background computed alongside is not kept as the primal: the primal
        stays the differentiable operator run that ``vjp`` needs."""
        with torch.no_grad():
            out = self._born(direction)
        return TangentResult(out[-1], TMState.from_tensors(tuple(out[7:14])))

    def second_vjp(self, direction: EMDirection, cotangent: torch.Tensor) -> EMGradient:
        """Gradient of ``<J(model) direction, cotangent>`` with respect to the
        linearization targets (reference ``linearization.py:129-146``): the
        Born receivers differentiated through the native adjoint of the Born
        recurrence (background and scattered adjoints coupled), the
        coefficient partials and the source terms by torch's chain rule."""
        out = self._born(direction)
        tangent = out[-1]
        if cotangent.shape != tangent.shape:
            raise ValueError("Receiver cotangent must match tangent receiver data shape.")
        grads = torch.autograd.grad(tangent, self._parameters(), cotangent, allow_unused=True)
        values = {n: None for n in _TARGETS}
        for n, g in zip((n for n in _TARGETS if n in self.targets), grads):
            values[n] = g
        return EMGradient(**values)

    def _born(self, direction: EMDirection):
        if self._closed:
            raise RuntimeError("The linearized Maxwell operator is closed.")
        direction.validate_for(self.model)
        if direction.mu is not None:
            raise NotImplementedError("TM2D tangent propagation for mu is not available.")
        from .tm2d_born import borntm

        op, model = self.operator, self.model
        d, e = op.discretization, op.experiment
        return borntm(
            model.epsilon, model.sigma, model.mu, grid_spacing=d.spacing, dt=d.dt,
            source_amplitude=e.source_amplitude,
            source_location=e.acquisition.source_location,
            receiver_location=e.acquisition.receiver_location,
            depsilon=direction.epsilon, dsigma=direction.sigma, stencil=d.stencil,
            pml_width=d.boundary.width, max_vel=d.max_velocity, nt=e.nt,
            linearize_source=e.source_convention is SourceConvention.PHYSICAL_CURRENT,
            freq_taper_frac=e.frequency_taper_fraction, time_pad_frac=e.time_padding_fraction,
            time_taper=e.time_taper,
        )

    def _separable_gradient(self, observed: torch.Tensor, residual):
        """Value and gradient of a shot-separable receiver objective, chunk
        by chunk with one FP64 accumulator (``tm2d.separable_gradient``).
        Also fills the primal (receivers and final state).  Returns None
        when the call needs the general path (structured batches, callbacks,
        host-staged storage, dispersion)."""
        op, model = self.operator, self.model
        d, e = op.discretization, op.experiment
        storage = self.storage
        if (model.epsilon.ndim != 2 or model.dispersion is not None
                or storage.mode.value not in ("device", "auto")
                or (storage.mode.value == "auto" and storage.bytes_limit_device is not None)
                or e.frequency_taper_fraction or e.time_padding_fraction or e.time_taper):
            return None
        from .tm2d import separable_gradient

        value, grads, r, fields, info = separable_gradient(
            model.epsilon, model.sigma, model.mu, d.spacing, d.dt, e.source_amplitude,
            e.acquisition.source_location, e.acquisition.receiver_location, observed,
            residual=residual, stencil=d.stencil, pml_width=d.boundary.width,
            max_vel=d.max_velocity, nt=e.nt,
            model_gradient_sampling_interval=op.model_gradient_sampling_interval,
            storage_compression=storage.compression,
            storage_bytes_limit_device=storage.bytes_limit_device,
            targets=tuple(n for n in ("epsilon", "sigma") if n in self.targets),
        )
        self.last_plan = info
        self._primal = ForwardResult(receiver_data=r, final_state=TMState.from_tensors(tuple(fields)))
        return value, EMGradient(epsilon=grads.get("epsilon"), sigma=grads.get("sigma"), mu=None)


class LinearizedMaxwell3D(_LinearizedMaxwell):
    def jvp(self, direction: EMDirection) -> TangentResult:
        """``J(model) direction``: receivers and final state of the 3-D Born
        scattered field (reference ``linearization.py:178-229`` over
        ``maxwell3d_born.py``).  ``second_vjp`` is not part of this build
        for 3-D."""
        if self._closed:
            raise RuntimeError("The linearized Maxwell operator is closed.")
        direction.validate_for(self.model)
        if direction.mu is not None:
            raise NotImplementedError("3-D tangent propagation for mu is not available.")
        from .maxwell3d import born3d

        op, model = self.operator, self.model
        d, e = op.discretization, op.experiment
        out = born3d(
            model.epsilon, model.sigma, model.mu, d.spacing, d.dt, e.source_amplitude,
            e.acquisition.source_location, e.acquisition.receiver_location,
            depsilon=direction.epsilon, dsigma=direction.sigma, stencil=d.stencil,
            pml_width=d.boundary.width, max_vel=d.max_velocity, nt=e.nt,
            source_component=e.source_component, receiver_component=e.receiver_component,
            linearize_source=e.source_convention is SourceConvention.PHYSICAL_CURRENT,
        )
        return TangentResult(out[-1], EM3DState.from_tensors(tuple(out[:18])))

    def _separable_gradient(self, observed: torch.Tensor, residual):
        """3-D counterpart of :meth:`LinearizedMaxwellTM._separable_gradient`
        (``maxwell3d.separable_gradient3d``)."""
        op, model = self.operator, self.model
        d, e = op.discretization, op.experiment
        storage = self.storage
        if (model.epsilon.ndim != 3 or model.dispersion is not None
                or storage.mode.value not in ("device", "auto")
                or (storage.mode.value == "auto" and storage.bytes_limit_device is not None)
                or e.frequency_taper_fraction or e.time_padding_fraction or e.time_taper):
            return None
        from .maxwell3d import separable_gradient3d

        value, grads, r, fields, info = separable_gradient3d(
            model.epsilon, model.sigma, model.mu, d.spacing, d.dt, e.source_amplitude,
            e.acquisition.source_location, e.acquisition.receiver_location, observed,
            residual=residual, stencil=d.stencil, pml_width=d.boundary.width,
            max_vel=d.max_velocity, nt=e.nt,
            model_gradient_sampling_interval=op.model_gradient_sampling_interval,
            source_component=e.source_component, receiver_component=e.receiver_component,
            storage_compression=storage.compression,
            storage_bytes_limit_device=storage.bytes_limit_device,
            storage_chunk_steps=storage.chunk_steps,
            targets=tuple(n for n in ("epsilon", "sigma") if n in self.targets),
        )
        self.last_plan = info
        self._primal = ForwardResult(receiver_data=r,
                                     final_state=EM3DState.from_tensors(tuple(fields)))
        return value, EMGradient(epsilon=grads.get("epsilon"), sigma=grads.get("sigma"), mu=None)


__all__ = ["LinearizedMaxwell3D", "LinearizedMaxwellTM"]
