"""Wavefield callbacks (reference ``src/tide/callbacks.py``).

A propagation called with ``forward_callback`` / ``backward_callback`` runs
its time loop in chunks of ``callback_frequency`` steps and hands the
callback a :class:`CallbackState` after each chunk: the padded wavefields
at that step, the models and, in the backward pass, the gradients
accumulated so far.  ``get_*`` return views of three regions:

* ``"full"``  -- the whole padded grid (CPML and the stencil halo);
* ``"pml"``   -- the model plus the CPML layers (halo removed);
* ``"inner"`` -- the physical model only.

Padding lists are ``[d0_low, d0_high, d1_low, d1_high(, d2_low, d2_high)]``
for the stencil halo (``fd_pad``) and the CPML (``pml_width``).
"""

from __future__ import annotations

from collections.abc import Callable, Sequence

import torch

_VIEWS = ("full", "pml", "inner")


class CallbackState:
    """Propagation state at one step, as seen by a user callback.

    Attributes: ``dt`` (s), ``step`` (the step just completed, in the
    user's time axis), ``nt`` (steps of the whole propagation) and
    ``is_backward``.
    """

    def __init__(
        self,
        dt: float,
        step: int,
        nt: int,
        wavefields: dict[str, torch.Tensor],
        models: dict[str, torch.Tensor],
        gradients: dict[str, torch.Tensor] | None = None,
        fd_pad: Sequence[int] | None = None,
        pml_width: Sequence[int] | None = None,
        is_backward: bool = False,
        grid_spacing: Sequence[float] | None = None,
    ) -> None:
        self.dt = dt
        self.step = step
        self.nt = nt
        self.is_backward = is_backward
        self._wavefields = dict(wavefields)
        self._models = dict(models)
        self._gradients = dict(gradients) if gradients else {}
        self._grid_spacing = grid_spacing
        self._ndim = self._spatial_ndim(fd_pad, pml_width, models)
        zeros = [0] * (2 * self._ndim)
        self._fd_pad = list(fd_pad) if fd_pad is not None else zeros
        self._pml_width = list(pml_width) if pml_width is not None else zeros

    @staticmethod
    def _spatial_ndim(fd_pad, pml_width, models) -> int:
        for pad in (fd_pad, pml_width):
            if pad is not None and len(pad) in (4, 6):
                return len(pad) // 2
        for value in models.values():
            if isinstance(value, torch.Tensor):
                # [ny,nx] / [N,ny,nx] are 2-D; [N,nz,ny,nx] is 3-D.  An
                # unbatched 3-D model needs the padding lists to say so.
                return 3 if value.ndim == 4 else 2
        return 2

    # ------------------------------------------------------------ summary
    @property
    def time(self) -> float:
        return self.step * self.dt

    @property
    def progress(self) -> float:
        return self.step / max(self.nt - 1, 1)

    @property
    def wavefield_names(self) -> list[str]:
        return list(self._wavefields)

    @property
    def model_names(self) -> list[str]:
        return list(self._models)

    @property
    def gradient_names(self) -> list[str]:
        return list(self._gradients)

    # ------------------------------------------------------------ access
    def get_wavefield(self, name: str, view: str = "inner") -> torch.Tensor:
        if name not in self._wavefields:
            raise KeyError(f"Wavefield '{name}' not found. Available: "
                           f"{', '.join(self._wavefields)}")
        return self._region(self._wavefields[name], view)

    def get_model(self, name: str, view: str = "inner") -> torch.Tensor:
        if name not in self._models:
            raise KeyError(f"Model '{name}' not found. Available: {', '.join(self._models)}")
        return self._region(self._models[name], view)

    def get_gradient(self, name: str, view: str = "inner") -> torch.Tensor:
        if not self._gradients:
            raise RuntimeError(
                "Gradients are only available during backward propagation. "
                "Use backward_callback instead of forward_callback."
            )
        if name not in self._gradients:
            raise KeyError(f"Gradient '{name}' not found. Available: "
                           f"{', '.join(self._gradients)}")
        return self._region(self._gradients[name], view)

    def _region(self, x, view: str):
        if not isinstance(x, torch.Tensor) or view == "full":
            if view not in _VIEWS:
                raise ValueError(f"view must be 'full', 'pml', or 'inner', but got '{view}'")
            return x
        if view not in _VIEWS:
            raise ValueError(f"view must be 'full', 'pml', or 'inner', but got '{view}'")
        nd = self._ndim
        lo = list(self._fd_pad[0::2][:nd])
        hi = list(self._fd_pad[1::2][:nd])
        if view == "inner":
            lo = [a + b for a, b in zip(lo, self._pml_width[0::2][:nd])]
            hi = [a + b for a, b in zip(hi, self._pml_width[1::2][:nd])]
        cut = tuple(slice(a, (x.shape[x.ndim - nd + i] - b) if b > 0 else None)
                    for i, (a, b) in enumerate(zip(lo, hi)))
        return x[(Ellipsis, *cut)]

    def __repr__(self) -> str:
        return (f"CallbackState(step={self.step}/{self.nt}, time={self.time:.2e}s, "
                f"is_backward={self.is_backward}, wavefields={self.wavefield_names}, "
                f"models={self.model_names})")


Callback = Callable[[CallbackState], None]

__all__ = ["Callback", "CallbackState"]
