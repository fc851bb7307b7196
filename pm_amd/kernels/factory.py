"""Backend registry: the reference's factory (pstm/kernels/factory.py:32-216)
with ``hip`` registered.

Only the HIP backend lives in this package; AUTO selects it when the library
is built and a GPU is visible (it would sit first in the reference's AUTO
priority list, factory.py:203-209).  There is deliberately no CPU backend
here: the CPU oracle under oracle/ is test infrastructure.
"""

from __future__ import annotations

from pm_amd.config import ComputeBackend, parse_backend
from pm_amd.kernels.base import MigrationKernel

_BACKEND_REGISTRY: dict[ComputeBackend, type] = {}

#: AUTO priority (GPU backends before CPU ones)
PRIORITY = [ComputeBackend.HIP, ComputeBackend.METAL_COMPILED, ComputeBackend.METAL_CPP,
            ComputeBackend.NUMBA_CPU, ComputeBackend.MLX_METAL, ComputeBackend.NUMPY]


def hip_available() -> bool:
    """Library built and a device visible (counting devices does not
    initialise HIP)."""
    from pm_amd.kernels import _abi

    if not _abi.library_path().exists():
        return False
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


def _register_backends() -> None:
    from pm_amd.kernels.hip_backend import HipKernel

    _BACKEND_REGISTRY[ComputeBackend.HIP] = HipKernel


def register_backend(backend: ComputeBackend, cls: type) -> None:
    """Add a backend class (the hook a reference install would use)."""
    _BACKEND_REGISTRY[backend] = cls


def get_available_backends() -> list[ComputeBackend]:
    if not _BACKEND_REGISTRY:
        _register_backends()
    out = []
    for b in _BACKEND_REGISTRY:
        if b == ComputeBackend.HIP and not hip_available():
            continue
        out.append(b)
    return out


def is_backend_available(backend: ComputeBackend) -> bool:
    return backend in get_available_backends()


def select_best_backend() -> ComputeBackend:
    available = get_available_backends()
    for b in PRIORITY:
        if b in available:
            return b
    raise RuntimeError("No compute backends available")


def create_kernel(backend: ComputeBackend | str) -> MigrationKernel:
    """Instantiate a backend; unknown or unavailable backends raise ValueError
    (factory.py:136-180)."""
    if not _BACKEND_REGISTRY:
        _register_backends()
    backend = parse_backend(backend)
    if backend == ComputeBackend.AUTO:
        backend = select_best_backend()
    if backend not in _BACKEND_REGISTRY:
        raise ValueError(f"Backend '{backend.value}' not available. "
                         f"Available: {[b.value for b in get_available_backends()]}")
    return _BACKEND_REGISTRY[backend]()


_register_backends()
