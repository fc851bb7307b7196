"""Backend enum, backend-string map and kernel defaults.

Mirrors ``ComputeBackend`` (pstm/config/models.py:72-79), ``MigrationKernelType``
(models.py:82-86), ``BACKEND_STRING_MAP``/``parse_backend``
(pstm/config/backends.py:13-54) with the new ``hip`` value added, and the
kernel defaults of the reference's settings (pstm/settings.py:90-135).
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum


class ComputeBackend(str, Enum):
    """Compute kernel backend (the reference's values plus ``hip``)."""

    AUTO = "auto"
    NUMPY = "numpy"
    NUMBA_CPU = "numba_cpu"
    MLX_METAL = "mlx_metal"
    METAL_CPP = "metal_cpp"
    METAL_COMPILED = "metal_compiled"
    HIP = "hip"


class MigrationKernelType(str, Enum):
    """Traveltime computation (models.py:82-86)."""

    STRAIGHT_RAY = "straight_ray"
    CURVED_RAY = "curved_ray"
    ANISOTROPIC_VTI = "anisotropic_vti"


BACKEND_STRING_MAP: dict[str, ComputeBackend] = {b.value: b for b in ComputeBackend}
VALID_BACKEND_STRINGS = list(BACKEND_STRING_MAP)


def parse_backend(backend: str | ComputeBackend) -> ComputeBackend:
    """String or enum -> ComputeBackend; unknown strings raise ValueError and
    other types TypeError, as backends.py:27-54 does."""
    if isinstance(backend, ComputeBackend):
        return backend
    if not isinstance(backend, str):
        raise TypeError(
            f"Backend must be string or ComputeBackend, got {type(backend).__name__}")
    key = backend.lower().strip()
    if key not in BACKEND_STRING_MAP:
        raise ValueError(f"Unknown backend '{backend}'. "
                         f"Valid backends: {', '.join(VALID_BACKEND_STRINGS)}")
    return BACKEND_STRING_MAP[key]


@dataclass(frozen=True)
class KernelDefaults:
    """Built-in values the reference's KernelConfig falls back to
    (settings.py:94-132; applied in base.py:239-259)."""

    max_aperture_m: float = 5000.0
    min_aperture_m: float = 500.0
    max_dip_degrees: float = 45.0
    taper_fraction: float = 0.1
    enable_antialiasing: bool = True
    apply_spreading_correction: bool = True
    default_interpolation: str = "linear"
    output_dt_ms: float = 2.0


DEFAULTS = KernelDefaults()
