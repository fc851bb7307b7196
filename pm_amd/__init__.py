"""pm_amd — MI355X-native `hip` backend for PSTM tile migration.

The package mirrors the reference's kernel boundary (pstm/kernels/base.py,
pstm/kernels/factory.py) and registers a ``hip`` backend whose compute runs in
hand-written gfx950 HIP kernels behind the C ABI of include/pstm_hip.h.
"""

from pm_amd.kernels import (  # noqa: F401
    KernelConfig,
    KernelMetrics,
    MigrationKernel,
    OutputTile,
    TraceBlock,
    VelocitySlice,
    create_kernel,
    create_output_tile,
    create_trace_block,
)

__version__ = "0.1.0"
