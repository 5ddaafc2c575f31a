"""Material coefficients and CPML profiles.

The formulas are those of the reference so that every coefficient the
kernels read is the same number (reference ``src/tide/utils.py``):
  * ``ca = (1 - s)/(1 + s)``, ``cb = (dt/eps)/(1 + s)`` with
    ``s = sigma dt / (2 eps)``, ``cq = dt/mu`` (``utils.py:13-41``);
  * CPML profiles with N = 4 grading, kappa_max = 5, alpha_max = 0.008,
    ``sigma0 = (N+1)/(150 pi dx)``, ``b = exp(-(sigma/kappa + alpha) dt/eps0)``
    and ``a = sigma (b-1) / (kappa (sigma + kappa alpha) + 1e-9)`` where
    sigma > 1e-6 (``utils.py:220-389``).
The torch operations are kept in the same order and dtype, so results are
bit-identical to the reference's on the same device.
"""

from __future__ import annotations

import math
from collections.abc import Sequence

import torch

EP0 = 8.8541878128e-12  # vacuum permittivity (F/m)
MU0 = 1.2566370614359173e-06  # vacuum permeability (H/m)
C0 = 1.0 / math.sqrt(EP0 * MU0)  # speed of light in vacuum (m/s)

_KAPPA_MAX = 5.0
_ALPHA_MAX = 0.008


def prepare_parameters(
    epsilon_r: torch.Tensor, sigma: torch.Tensor, mu_r: torch.Tensor, dt: float
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Update coefficients (ca, cb, cq) from relative permittivity,
    conductivity (S/m) and relative permeability."""
    eps_abs = epsilon_r * EP0
    mu_abs = mu_r * MU0
    loss = sigma * dt / (2.0 * eps_abs)
    denom = 1.0 + loss
    ca = (1.0 - loss) / denom
    cb = (dt / eps_abs) / denom
    cq = dt / mu_abs
    return ca, cb, cq


def debye_parameter(value, *, reference: torch.Tensor, name: str) -> torch.Tensor:
    """A Debye parameter as ``[n_poles, *model_shape]``: scalars and
    model-shaped (broadcastable) tensors are one pole, a leading extra axis
    counts the poles (reference ``utils.py:44-76``)."""
    t = torch.as_tensor(value, device=reference.device, dtype=reference.dtype)
    shape = tuple(reference.shape)
    if t.ndim <= len(shape):
        try:
            torch.broadcast_shapes(tuple(t.shape), shape)
        except RuntimeError as exc:
            raise ValueError(f"{name} with shape {tuple(t.shape)} is not broadcastable "
                             f"to model shape {shape}.") from exc
        t = t.reshape((1,) + tuple(t.shape))
    elif t.ndim == len(shape) + 1:
        try:
            torch.broadcast_shapes(tuple(t.shape[1:]), shape)
        except RuntimeError as exc:
            raise ValueError(f"{name} with shape {tuple(t.shape)} is not broadcastable "
                             f"to model shape {shape} on trailing dimensions.") from exc
    else:
        raise ValueError(f"{name} must be a scalar, model-shaped tensor, or "
                         f"[n_poles, *model_shape], but got shape {tuple(t.shape)}.")
    return torch.broadcast_to(t, (t.shape[0], *shape))


def validate_material_inputs(epsilon_r: torch.Tensor, *, dispersion=None,
                             dt: float | None = None) -> None:
    """Positive permittivity; for Debye poles delta_epsilon >= 0, tau > 0 and
    dt < min(tau) (reference ``utils.py:79-116``)."""
    if torch.any(epsilon_r <= 0):
        raise ValueError("epsilon must be strictly positive.")
    if dispersion is None:
        return
    if dt is None:
        raise ValueError("dt must be provided when validating dispersion inputs.")
    d_eps = debye_parameter(dispersion.delta_epsilon, reference=epsilon_r, name="delta_epsilon")
    tau = debye_parameter(dispersion.tau, reference=epsilon_r, name="tau")
    if torch.any(d_eps < 0):
        raise ValueError("delta_epsilon must be non-negative.")
    if torch.any(tau <= 0):
        raise ValueError("tau must be strictly positive.")
    min_tau = float(tau.detach().amin().item())
    if dt >= min_tau:
        raise ValueError(
            f"Debye dispersion requires dt < min(tau), but got dt={dt} and min(tau)={min_tau}.")


def compile_material_coefficients(
    epsilon_r: torch.Tensor,
    sigma: torch.Tensor,
    mu_r: torch.Tensor,
    dt: float,
    *,
    dispersion=None,
    validate_inputs: bool = True,
) -> dict:
    """Solver coefficients (reference ``utils.py:119-186``).

    Without dispersion: ``ca, cb, cq`` of the plain update.  With Debye
    poles (``dispersion.delta_epsilon``, ``dispersion.tau``) the
    polarization of pole p obeys the trapezoidal recurrence
    ``P_p = a_p P_p + b_p (E^{n+1} + E^n)``, ``a_p = (2 tau_p - dt) /
    (2 tau_p + dt)``, ``b_p = EP0 delta_eps_p dt / (2 tau_p + dt)``, and the
    E update becomes ``E^{n+1} = ca E^n + cb curl H + sum_p cp_p P_p`` with
    ``den = EP0 eps_inf + sigma dt / 2 + sum_p b_p``, ``ca = (EP0 eps_inf -
    sigma dt / 2 - sum_p b_p) / den``, ``cb = dt / den``, ``cp_p = (1 - a_p) /
    den``; the result also holds ``has_dispersion`` and ``debye`` (``a``,
    ``b``, ``cp`` as ``[n_poles, *model_shape]``, ``n_poles``)."""
    if validate_inputs:
        validate_material_inputs(epsilon_r, dispersion=dispersion, dt=dt)
    ca, cb, cq = prepare_parameters(epsilon_r, sigma, mu_r, dt)
    out = {"ca": ca, "cb": cb, "cq": cq, "dispersion": dispersion, "has_dispersion": False}
    if dispersion is None:
        return out
    d_eps = debye_parameter(dispersion.delta_epsilon, reference=epsilon_r, name="delta_epsilon")
    tau = debye_parameter(dispersion.tau, reference=epsilon_r, name="tau")
    eps_inf = epsilon_r * EP0
    tau_term = 2.0 * tau + dt
    a = (2.0 * tau - dt) / tau_term
    b = (EP0 * d_eps * dt) / tau_term
    b_sum = b.sum(dim=0)
    den = eps_inf + sigma * dt / 2.0 + b_sum
    out.update({
        "ca": (eps_inf - sigma * dt / 2.0 - b_sum) / den,
        "cb": dt / den,
        "cq": dt / (mu_r * MU0),
        "has_dispersion": True,
        "debye": {"a": a, "b": b, "cp": (1.0 - a) / den.unsqueeze(0), "n_poles": int(a.shape[0])},
    })
    return out


def coefficient_partials(
    ca: torch.Tensor, cb: torch.Tensor, dt: float
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Closed-form partials of (ca, cb) w.r.t. (epsilon_r, sigma):
    ``dca/deps_r = EP0 (1-ca) cb/dt``, ``dcb/deps_r = -EP0 cb^2/dt``,
    ``dca/dsigma = -(1+ca) cb/2``, ``dcb/dsigma = -cb^2/2``."""
    cb2 = cb * cb
    return (
        EP0 * (1.0 - ca) * cb / dt,
        -EP0 * cb2 / dt,
        -0.5 * (1.0 + ca) * cb,
        -0.5 * cb2,
    )


def setup_pml(
    pml_width: Sequence[int],
    pml_start: Sequence[float],
    max_pml: float,
    dt: float,
    n: int,
    max_vel: float,
    dtype: torch.dtype,
    device: torch.device,
    pml_freq: float,
    start: float = 0.0,
    r_val: float = 1e-8,
    n_power: int = 4,
    eps: float = 1e-9,
    grid_spacing: float = 1.0,
    eps_scale: float = EP0,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """CPML (a, b, kappa) along one axis of length ``n``.

    ``pml_width`` = [low, high] cells, ``pml_start`` = inner edge coordinate
    of each strip (in cells), ``start`` = coordinate of element 0 (0 for
    integer points, 0.5 for half points).  ``max_vel``, ``pml_freq`` and
    ``r_val`` are accepted for signature compatibility; the electromagnetic
    grading does not use them.
    """
    del max_vel, pml_freq, r_val
    a = torch.zeros(n, device=device, dtype=dtype)
    b = torch.zeros(n, device=device, dtype=dtype)
    k = torch.ones(n, device=device, dtype=dtype)
    if max_pml == 0 or (pml_width[0] == 0 and pml_width[1] == 0):
        return a, b, k

    sigma0 = (n_power + 1) / (150.0 * math.pi * grid_spacing)
    coord = torch.arange(start, start + n, device=device, dtype=dtype)

    for side in (0, 1):
        width = pml_width[side]
        if width <= 0:
            continue
        depth = pml_start[0] - coord if side == 0 else coord - pml_start[1]
        inside = depth >= 0
        frac = torch.clamp(depth / width, 0, 1)
        graded = frac**n_power
        sig = sigma0 * graded
        kap = 1.0 + (_KAPPA_MAX - 1.0) * graded
        alpha = _ALPHA_MAX * (1.0 - frac) + 0.1 * _ALPHA_MAX
        k = torch.where(inside, kap, k)
        b_side = torch.exp(-(sig / kap + alpha) * dt / eps_scale)
        b = torch.where(inside, b_side, b)
        a_side = sig * (b_side - 1.0) / (kap * (sig + kap * alpha) + eps)
        a_side = torch.where(sig > 1e-6, a_side, torch.zeros_like(a_side))
        a = torch.where(inside, a_side, a)
    return a, b, k


def setup_pml_half(
    pml_width: Sequence[int],
    pml_start: Sequence[float],
    max_pml: float,
    dt: float,
    n: int,
    max_vel: float,
    dtype: torch.dtype,
    device: torch.device,
    pml_freq: float,
    start: float = 0.0,
    r_val: float = 1e-8,
    n_power: int = 4,
    eps: float = 1e-9,
    grid_spacing: float = 1.0,
    eps_scale: float = EP0,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """CPML profiles at half-integer points (coordinate offset +0.5)."""
    return setup_pml(
        pml_width,
        pml_start,
        max_pml,
        dt,
        n,
        max_vel,
        dtype,
        device,
        pml_freq,
        start=start + 0.5,
        r_val=r_val,
        n_power=n_power,
        eps=eps,
        grid_spacing=grid_spacing,
        eps_scale=eps_scale,
    )


__all__ = [
    "C0",
    "EP0",
    "MU0",
    "coefficient_partials",
    "compile_material_coefficients",
    "prepare_parameters",
    "setup_pml",
    "setup_pml_half",
    "validate_material_inputs",
]
