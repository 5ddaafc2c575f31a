from .types import (
    BackendPreference,
    Dimension,
    FallbackPolicy,
    StorageMode,
    StorageOptions,
)

__all__ = [
    "BackendPreference",
    "Dimension",
    "FallbackPolicy",
    "StorageMode",
    "StorageOptions",
]
