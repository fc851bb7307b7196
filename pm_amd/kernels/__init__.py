"""Kernel boundary and the ``hip`` backend (mirrors pstm/kernels/__init__.py)."""

from pm_amd.kernels.base import (  # noqa: F401
    KernelCapability,
    KernelConfig,
    KernelMetrics,
    MigrationKernel,
    OutputTile,
    TraceBlock,
    VelocitySlice,
    create_output_tile,
    create_trace_block,
)
from pm_amd.kernels.factory import (  # noqa: F401
    create_kernel,
    get_available_backends,
    is_backend_available,
    register_backend,
    select_best_backend,
)
from pm_amd.kernels.hip_backend import DeviceTraces, HipKernel  # noqa: F401
