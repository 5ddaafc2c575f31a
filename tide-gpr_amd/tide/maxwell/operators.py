"""Nonlinear Maxwell operators (reference ``src/tide/maxwell/operators.py``).

``MaxwellTM(discretization, experiment, execution=, storage=,
model_gradient_sampling_interval=)`` called on an :class:`EMModel` returns a
:class:`ForwardResult` holding receivers ``[nt, shots, n_rec]`` and the final
:class:`TMState`; ``linearize`` returns the adjoint-Jacobian view.
"""

from __future__ import annotations

from collections.abc import Iterable

from ..core import StorageOptions
from .contracts import Discretization, EMModel, ExecutionOptions, Experiment, Observers
from .results import EM3DState, ForwardResult, TMState


class MaxwellTM:
    """Differentiable 2-D transverse-magnetic Maxwell operator."""

    def __init__(
        self,
        discretization: Discretization,
        experiment: Experiment,
        *,
        execution: ExecutionOptions = ExecutionOptions(),
        storage: StorageOptions = StorageOptions(),
        model_gradient_sampling_interval: int = 1,
    ) -> None:
        if experiment.acquisition.spatial_ndim not in (None, 2):
            raise ValueError("MaxwellTM requires two-dimensional acquisition coordinates.")
        self.discretization = discretization
        self.experiment = experiment
        self.execution = execution
        self.storage = storage
        self.model_gradient_sampling_interval = model_gradient_sampling_interval

    def forward(
        self,
        model: EMModel,
        *,
        state: TMState | None = None,
        observers: Observers = Observers(),
    ) -> ForwardResult[TMState]:
        if model.epsilon.ndim not in (2, 3):
            raise ValueError("MaxwellTM model tensors must be 2-D or batched 3-D tensors.")
        from .tm2d import maxwelltm

        init = (None,) * 7 if state is None else (
            state.Ey, state.Hx, state.Hz, state.m_Ey_x, state.m_Ey_z, state.m_Hx_z, state.m_Hz_x
        )
        d, e = self.discretization, self.experiment
        out = maxwelltm(
            model.epsilon, model.sigma, model.mu,
            grid_spacing=d.spacing, dt=d.dt,
            source_amplitude=e.source_amplitude,
            source_location=e.acquisition.source_location,
            receiver_location=e.acquisition.receiver_location,
            stencil=d.stencil, pml_width=d.boundary.width, max_vel=d.max_velocity,
            Ey_0=init[0], Hx_0=init[1], Hz_0=init[2],
            m_Ey_x=init[3], m_Ey_z=init[4], m_Hx_z=init[5], m_Hz_x=init[6],
            nt=e.nt,
            model_gradient_sampling_interval=self.model_gradient_sampling_interval,
            freq_taper_frac=e.frequency_taper_fraction,
            time_pad_frac=e.time_padding_fraction,
            time_taper=e.time_taper,
            forward_callback=observers.forward,
            backward_callback=observers.backward,
            callback_frequency=observers.frequency,
            python_backend=self.execution.legacy_backend_request,
            storage_mode=self.storage.mode.value,
            storage_path=self.storage.path,
            storage_compression=self.storage.compression,
            storage_bytes_limit_device=self.storage.bytes_limit_device,
            storage_bytes_limit_host=self.storage.bytes_limit_host,
            storage_chunk_steps=self.storage.chunk_steps,
            n_threads=self.execution.n_threads,
            dispersion=model.dispersion,
            fallback=self.execution.fallback.value,
        )
        return ForwardResult(receiver_data=out[-1], final_state=TMState.from_tensors(tuple(out[:-1])))

    __call__ = forward

    def linearize(
        self,
        model: EMModel,
        *,
        storage: StorageOptions | None = None,
        targets: Iterable[str] = ("epsilon", "sigma"),
    ):
        from .linearization import LinearizedMaxwellTM

        return LinearizedMaxwellTM(self, model, storage=storage or self.storage, targets=targets)


class Maxwell3D:
    """Differentiable full 3-D Maxwell operator."""

    def __init__(
        self,
        discretization: Discretization,
        experiment: Experiment,
        *,
        execution: ExecutionOptions = ExecutionOptions(),
        storage: StorageOptions = StorageOptions(),
        model_gradient_sampling_interval: int = 1,
    ) -> None:
        if experiment.acquisition.spatial_ndim not in (None, 3):
            raise ValueError("Maxwell3D requires three-dimensional acquisition coordinates.")
        self.discretization = discretization
        self.experiment = experiment
        self.execution = execution
        self.storage = storage
        self.model_gradient_sampling_interval = model_gradient_sampling_interval

    def forward(
        self,
        model: EMModel,
        *,
        state: EM3DState | None = None,
        observers: Observers = Observers(),
    ) -> ForwardResult[EM3DState]:
        if model.epsilon.ndim not in (3, 4):
            raise ValueError("Maxwell3D model tensors must be 3-D or batched 4-D tensors.")
        from .maxwell3d import maxwell3d

        d, e = self.discretization, self.experiment
        init = None if state is None else tuple(
            getattr(state, n) for n in EM3DState.__dataclass_fields__
        )
        out = maxwell3d(
            model.epsilon, model.sigma, model.mu,
            grid_spacing=d.spacing, dt=d.dt,
            source_amplitude=e.source_amplitude,
            source_location=e.acquisition.source_location,
            receiver_location=e.acquisition.receiver_location,
            stencil=d.stencil, pml_width=d.boundary.width, max_vel=d.max_velocity,
            initial=init, nt=e.nt,
            model_gradient_sampling_interval=self.model_gradient_sampling_interval,
            forward_callback=observers.forward,
            backward_callback=observers.backward,
            source_component=e.source_component,
            receiver_component=e.receiver_component,
            python_backend=self.execution.legacy_backend_request,
            storage_mode=self.storage.mode.value,
            storage_compression=self.storage.compression,
            storage_bytes_limit_device=self.storage.bytes_limit_device,
            storage_bytes_limit_host=self.storage.bytes_limit_host,
            storage_chunk_steps=self.storage.chunk_steps,
            dispersion=model.dispersion,
        )
        return ForwardResult(receiver_data=out[-1], final_state=EM3DState.from_tensors(tuple(out[:-1])))

    __call__ = forward

    def linearize(
        self,
        model: EMModel,
        *,
        storage: StorageOptions | None = None,
        targets: Iterable[str] = ("epsilon", "sigma"),
    ):
        from .linearization import LinearizedMaxwell3D

        return LinearizedMaxwell3D(self, model, storage=storage or self.storage, targets=targets)


__all__ = ["Maxwell3D", "MaxwellTM"]
