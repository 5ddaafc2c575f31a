"""Source wavelets (reference ``src/tide/wavelets.py``)."""

from __future__ import annotations

import math
import operator

import torch


def _positive(name: str, value: float) -> float:
    try:
        out = float(value)
    except (TypeError, ValueError) as exc:
        raise TypeError(f"{name} must be a finite positive scalar.") from exc
    if not math.isfinite(out) or out <= 0.0:
        raise ValueError(f"{name} must be a finite positive scalar.")
    return out


def _time_axis(freq, length, dt, peak_time, dtype, device):
    f = _positive("freq", freq)
    step = _positive("dt", dt)
    if isinstance(length, bool):
        raise TypeError("length must be a positive integer.")
    n = operator.index(length)
    if n <= 0:
        raise ValueError("length must be a positive integer.")
    centre = 1.0 / f if peak_time is None else float(peak_time)
    if not math.isfinite(centre):
        raise ValueError("peak_time must be a finite scalar.")
    if dtype is not None and not dtype.is_floating_point:
        raise TypeError("dtype must be a floating-point torch dtype.")
    t = torch.arange(n, dtype=torch.get_default_dtype() if dtype is None else dtype, device=device)
    return t * step - centre, f


def ricker(freq, length, dt, peak_time=None, dtype=None, device=None) -> torch.Tensor:
    """Zero-phase Ricker wavelet with unit peak, peak at ``1/freq`` by default:
    ``(1 - 2 (pi f t)^2) exp(-(pi f t)^2)``."""
    t, f = _time_axis(freq, length, dt, peak_time, dtype, device)
    p2 = (math.pi * f * t).square()
    return (1.0 - 2.0 * p2) * torch.exp(-p2)


def gaussian(freq, length, dt, peak_time=None, dtype=None, device=None) -> torch.Tensor:
    """Unit-amplitude Gaussian pulse ``exp(-(pi f t)^2)``."""
    t, f = _time_axis(freq, length, dt, peak_time, dtype, device)
    return torch.exp(-(math.pi * f * t).square())


def gaussian_derivative(freq, length, dt, peak_time=None, dtype=None, device=None) -> torch.Tensor:
    """First-derivative Gaussian normalised to unit peak magnitude."""
    t, f = _time_axis(freq, length, dt, peak_time, dtype, device)
    ph = math.pi * f * t
    return -math.sqrt(2.0 * math.e) * ph * torch.exp(-ph.square())


__all__ = ["gaussian", "gaussian_derivative", "ricker"]
