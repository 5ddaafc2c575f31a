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


def validate_material_inputs(epsilon_r: torch.Tensor) -> None:
    if torch.any(epsilon_r <= 0):
        raise ValueError("epsilon must be strictly positive.")


def compile_material_coefficients(
    epsilon_r: torch.Tensor,
    sigma: torch.Tensor,
    mu_r: torch.Tensor,
    dt: float,
    *,
    validate_inputs: bool = True,
) -> dict[str, torch.Tensor]:
    """Coefficients for the non-dispersive update."""
    if validate_inputs:
        validate_material_inputs(epsilon_r)
    ca, cb, cq = prepare_parameters(epsilon_r, sigma, mu_r, dt)
    return {"ca": ca, "cb": cb, "cq": cq}


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
