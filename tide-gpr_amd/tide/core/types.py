"""Policy value objects shared by the public API (reference
``src/tide/core/types.py``).  Python 3.10 has no ``enum.StrEnum``; the
string enums here subclass ``str`` and ``Enum`` so they compare and format
as their values, as the reference's do."""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum


class _StrEnum(str, Enum):
    def __str__(self) -> str:
        return str(self.value)


class Dimension(_StrEnum):
    TM2D = "tm2d"
    EM3D = "em3d"


class BackendPreference(_StrEnum):
    AUTO = "auto"
    REFERENCE = "reference"
    NATIVE = "native"


class FallbackPolicy(_StrEnum):
    REFERENCE = "reference"
    ERROR = "error"


class StorageMode(_StrEnum):
    AUTO = "auto"
    DEVICE = "device"
    CPU = "cpu"
    DISK = "disk"
    NONE = "none"


@dataclass(frozen=True)
class StorageOptions:
    """Snapshot storage policy."""

    mode: StorageMode = StorageMode.DEVICE
    path: str = "."
    compression: bool | str = False
    bytes_limit_device: int | None = None
    bytes_limit_host: int | None = None
    chunk_steps: int = 0

    def __post_init__(self) -> None:
        if not isinstance(self.mode, StorageMode):
            object.__setattr__(self, "mode", StorageMode(str(self.mode).lower()))
        if not isinstance(self.path, str) or not self.path:
            raise ValueError("storage path must be a non-empty string.")
        for name, value in (
            ("bytes_limit_device", self.bytes_limit_device),
            ("bytes_limit_host", self.bytes_limit_host),
        ):
            if value is not None and value < 0:
                raise ValueError(f"{name} must be non-negative when provided.")
        if self.chunk_steps < 0:
            raise ValueError("storage_chunk_steps must be non-negative.")


__all__ = [
    "BackendPreference",
    "Dimension",
    "FallbackPolicy",
    "StorageMode",
    "StorageOptions",
]
