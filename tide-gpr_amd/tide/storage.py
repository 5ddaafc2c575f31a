"""Snapshot storage policy (reference ``src/tide/storage.py``).

Mode and format codes are the reference's ABI values.  ``device`` keeps the
snapshots in HBM, ``none`` stores nothing; ``auto`` resolves to ``device``
when the estimate fits the device byte limit.  The host-staged modes
(``cpu``, ``disk``) are not part of this build and raise.
"""

from __future__ import annotations

import torch

STORAGE_DEVICE = 0
STORAGE_CPU = 1
STORAGE_DISK = 2
STORAGE_NONE = 3

STORAGE_FORMAT_FULL = 0
STORAGE_FORMAT_BF16 = 1


def _normalize_storage_compression(storage_compression: bool | str | None) -> str:
    if storage_compression is True:
        return "bf16"
    if storage_compression is False or storage_compression is None:
        return "none"
    if isinstance(storage_compression, str):
        value = storage_compression.strip().lower()
        if value in {"none", "false", "off", "0"}:
            return "none"
        if value in {"bf16", "bfloat16"}:
            return "bf16"
    raise ValueError("storage_compression must be False/True or one of 'none' or 'bf16'.")


def _resolve_storage_compression(
    storage_compression: bool | str | None,
    dtype: torch.dtype,
    device: torch.device,
    *,
    context: str,
) -> tuple[str, torch.dtype, int, int]:
    """-> (kind, payload dtype, bytes per element, format code)."""
    del device
    kind = _normalize_storage_compression(storage_compression)
    if kind == "none":
        return kind, dtype, dtype.itemsize, STORAGE_FORMAT_FULL
    if dtype != torch.float32:
        raise NotImplementedError(f"{context} (BF16 storage) is only supported for float32.")
    return kind, torch.bfloat16, 2, STORAGE_FORMAT_BF16


def storage_mode_to_int(mode: str) -> int:
    table = {"device": STORAGE_DEVICE, "cpu": STORAGE_CPU, "disk": STORAGE_DISK, "none": STORAGE_NONE}
    try:
        return table[mode.lower()]
    except KeyError:
        raise ValueError(
            "storage_mode must be 'device', 'cpu', 'disk', 'none', or 'auto', "
            f"but got {mode!r}"
        ) from None


def resolve_storage_mode(
    mode: str,
    *,
    needs_storage: bool,
    total_bytes: int,
    bytes_limit_device: int | None,
    bytes_limit_host: int | None,
) -> str:
    """Reference auto-policy (``maxwell/tm2d_cuda.py:433-482``)."""
    mode = mode.lower()
    if mode not in {"device", "cpu", "disk", "none", "auto"}:
        raise ValueError(
            "storage_mode must be 'device', 'cpu', 'disk', 'none', or 'auto', "
            f"but got {mode!r}"
        )
    if not needs_storage:
        return "none" if mode == "auto" else mode
    if mode == "none":
        raise ValueError(
            "storage_mode='none' is not compatible with gradient computation "
            "when model parameters require gradients."
        )
    if mode == "auto":
        limit_dev = float("inf") if bytes_limit_device is None else bytes_limit_device
        limit_host = float("inf") if bytes_limit_host is None else bytes_limit_host
        if total_bytes <= limit_dev:
            return "device"
        return "cpu" if total_bytes <= limit_host else "disk"
    return mode


__all__ = [
    "STORAGE_CPU",
    "STORAGE_DEVICE",
    "STORAGE_DISK",
    "STORAGE_FORMAT_BF16",
    "STORAGE_FORMAT_FULL",
    "STORAGE_NONE",
    "resolve_storage_mode",
    "storage_mode_to_int",
]
