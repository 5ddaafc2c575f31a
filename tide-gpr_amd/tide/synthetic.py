"""Seeded synthetic models and acquisitions for benchmarks and tests.

Layered background plus Gaussian anomalies, the generator the benchmark
workloads are specified with:
  * ``n_layers`` horizontal layers, integer thickness U[20, 150] drawn top
    down (the last layer extends to the bottom), eps_r per layer U[1, 25];
  * ``n_anomalies`` Gaussian anomalies ``A exp(-r^2 / 2R^2)``, centre
    uniform, radius U[5, 40] cells, amplitude U[-5, 5]; eps_r clamped >= 1;
  * sigma i.i.d. U[0, 0.01] S/m; mu_r = 1.
All draws use ``torch.Generator('cpu')`` with the given seed.

The 3-D generator (:func:`layered_model_3d`) layers along z in the same way
and adds ellipsoidal Gaussian anomalies ``A exp(-sum_i (d_i / R_i)^2 / 2)``
with a semi-axis U[5, 40] cells drawn per axis.
"""

from __future__ import annotations

import math

import torch

from .utils import C0


def layered_model(
    ny: int,
    nx: int,
    *,
    seed: int,
    n_layers: int = 8,
    n_anomalies: int = 12,
    eps_range: tuple[float, float] = (1.0, 25.0),
    sigma_max: float = 0.01,
    dtype: torch.dtype = torch.float32,
    device: torch.device | str = "cpu",
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    g = torch.Generator(device="cpu").manual_seed(seed)
    eps = torch.empty(ny, nx, dtype=torch.float64)
    top = 0
    for layer in range(n_layers):
        value = eps_range[0] + (eps_range[1] - eps_range[0]) * torch.rand(1, generator=g, dtype=torch.float64).item()
        thick = int(torch.randint(20, 151, (1,), generator=g).item())
        bottom = ny if layer == n_layers - 1 else min(ny, top + thick)
        eps[top:bottom] = value
        top = bottom
        if top >= ny:
            break
    if top < ny:
        eps[top:] = eps[top - 1] if top > 0 else eps_range[0]
    yy = torch.arange(ny, dtype=torch.float64)[:, None]
    xx = torch.arange(nx, dtype=torch.float64)[None, :]
    for _ in range(n_anomalies):
        cy = torch.rand(1, generator=g, dtype=torch.float64).item() * ny
        cx = torch.rand(1, generator=g, dtype=torch.float64).item() * nx
        radius = 5.0 + 35.0 * torch.rand(1, generator=g, dtype=torch.float64).item()
        amp = -5.0 + 10.0 * torch.rand(1, generator=g, dtype=torch.float64).item()
        eps += amp * torch.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2.0 * radius**2))
    eps.clamp_(min=1.0)
    sigma = sigma_max * torch.rand(ny, nx, generator=g, dtype=torch.float64)
    mu = torch.ones(ny, nx, dtype=torch.float64)
    return tuple(t.to(device=device, dtype=dtype) for t in (eps, sigma, mu))  # type: ignore[return-value]


def layered_model_3d(
    nz: int,
    ny: int,
    nx: int,
    *,
    seed: int,
    n_layers: int = 8,
    n_anomalies: int = 20,
    eps_range: tuple[float, float] = (1.0, 25.0),
    sigma_max: float = 0.01,
    dtype: torch.dtype = torch.float32,
    device: torch.device | str = "cpu",
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """3-D model [nz, ny, nx]: z layers of thickness U[20, 150] (eps_r
    U[1, 25] each), ellipsoidal anomalies (centre uniform, semi-axes U[5, 40]
    cells per axis, amplitude U[-5, 5]), eps_r clamped >= 1, sigma
    U[0, sigma_max], mu_r = 1."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    eps = torch.empty(nz, ny, nx, dtype=torch.float64)
    top = 0
    for layer in range(n_layers):
        value = eps_range[0] + (eps_range[1] - eps_range[0]) * torch.rand(1, generator=g, dtype=torch.float64).item()
        thick = int(torch.randint(20, 151, (1,), generator=g).item())
        bottom = nz if layer == n_layers - 1 else min(nz, top + thick)
        eps[top:bottom] = value
        top = bottom
        if top >= nz:
            break
    if top < nz:
        eps[top:] = eps[top - 1] if top > 0 else eps_range[0]
    axes = [torch.arange(n, dtype=torch.float64) for n in (nz, ny, nx)]
    for _ in range(n_anomalies):
        centre = [torch.rand(1, generator=g, dtype=torch.float64).item() * n for n in (nz, ny, nx)]
        radii = [5.0 + 35.0 * torch.rand(1, generator=g, dtype=torch.float64).item() for _ in range(3)]
        amp = -5.0 + 10.0 * torch.rand(1, generator=g, dtype=torch.float64).item()
        # separable: exp of a sum = product of per-axis factors
        fz, fy, fx = (torch.exp(-0.5 * ((a - c) / r) ** 2) for a, c, r in zip(axes, centre, radii))
        eps += amp * fz[:, None, None] * fy[None, :, None] * fx[None, None, :]
    eps.clamp_(min=1.0)
    sigma = sigma_max * torch.rand(nz, ny, nx, generator=g, dtype=torch.float64)
    mu = torch.ones(nz, ny, nx, dtype=torch.float64)
    return tuple(t.to(device=device, dtype=dtype) for t in (eps, sigma, mu))  # type: ignore[return-value]


def add_anomaly(eps: torch.Tensor, cy: float, cx: float, radius: float, amp: float) -> torch.Tensor:
    ny, nx = eps.shape[-2:]
    yy = torch.arange(ny, dtype=torch.float64, device=eps.device)[:, None]
    xx = torch.arange(nx, dtype=torch.float64, device=eps.device)[None, :]
    bump = amp * torch.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2.0 * radius**2))
    return (eps.double() + bump).clamp(min=1.0).to(eps.dtype)


def cfl_dt(spacing: float, ndim: int = 2, courant: float = 0.9, eps_min: float = 1.0) -> float:
    """``courant * dx / (sqrt(ndim) * c)`` with ``c = C0 / sqrt(eps_min)``."""
    return courant * spacing / (math.sqrt(ndim) * C0 / math.sqrt(eps_min))


def surface_acquisition(
    n_shots: int,
    n_receivers: int,
    *,
    src_spacing: int = 16,
    src_offset: int = 8,
    rec_spacing: int = 4,
    rec_offset: int = 2,
    row: int = 2,
    device: torch.device | str = "cpu",
) -> tuple[torch.Tensor, torch.Tensor]:
    """Sources at ``(row, src_offset + src_spacing k)``, receivers at
    ``(row, rec_offset + rec_spacing r)`` shared by every shot."""
    src = torch.tensor([[[row, src_offset + src_spacing * s]] for s in range(n_shots)], device=device)
    rec_one = torch.tensor([[row, rec_offset + rec_spacing * k] for k in range(n_receivers)], device=device)
    return src, rec_one[None].expand(n_shots, -1, -1).contiguous()


__all__ = ["add_anomaly", "cfl_dt", "layered_model", "layered_model_3d", "surface_acquisition"]
