"""Kernel boundary: the data structures and Protocol every backend implements.

Same interface as the reference's pstm/kernels/base.py so that a caller written
against it (the tiled executor, pstm/pipeline/executor.py:1620-1777; the tests,
tests/test_kernels.py) can hand its objects to the ``hip`` backend unchanged.

Differences, all additive:
  * array fields of TraceBlock / OutputTile / VelocitySlice may also be torch
    tensors already resident on the GPU (the HIP backend then skips the
    host->device copy);
  * KernelConfig has ``antialias`` ("none" | "triangle") and
    ``aa_max_half_width``: the reference's ``aa_enabled`` is ignored by the
    registered Numba kernel (numba_cpu_optimized.py:593-618 passes no AA
    argument), so the backend keeps that behaviour and only runs triangle
    anti-aliasing when ``antialias="triangle"`` is asked for.
"""

from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Any, Protocol, runtime_checkable

import numpy as np
from numpy.typing import NDArray

from pm_amd.config import DEFAULTS


class KernelCapability(Enum):
    """Capabilities a kernel may advertise (base.py:17-24)."""

    FP16 = "fp16"
    FP64 = "fp64"
    COMPLEX = "complex"
    ASYNC = "async"
    BATCH = "batch"


def _n(a: Any) -> int:
    return int(a.shape[0])


@dataclass
class TraceBlock:
    """Input traces and their geometry (base.py:27-116).

    amplitudes (n_traces, n_samples) float32, C-contiguous; coordinates
    (n_traces,) float64 absolute survey coordinates.
    """

    amplitudes: NDArray[np.float32]
    source_x: NDArray[np.float64]
    source_y: NDArray[np.float64]
    receiver_x: NDArray[np.float64]
    receiver_y: NDArray[np.float64]
    offset: NDArray[np.float64]
    midpoint_x: NDArray[np.float64]
    midpoint_y: NDArray[np.float64]
    weights: NDArray[np.float64] | None = None
    sample_rate_ms: float = 2.0
    start_time_ms: float = 0.0

    @property
    def n_traces(self) -> int:
        return _n(self.amplitudes)

    @property
    def n_samples(self) -> int:
        return int(self.amplitudes.shape[1])

    @property
    def time_axis_ms(self) -> NDArray[np.float64]:
        return np.arange(self.n_samples) * self.sample_rate_ms + self.start_time_ms

    def validate(self) -> list[str]:
        """Consistency errors, empty when valid (base.py:76-100)."""
        errors: list[str] = []
        n = self.n_traces
        for name in ("source_x", "source_y", "receiver_x", "receiver_y",
                     "offset", "midpoint_x", "midpoint_y"):
            length = _n(getattr(self, name))
            if length != n:
                errors.append(f"{name} length mismatch: {length} != {n}")
        amp = self.amplitudes
        if isinstance(amp, np.ndarray):
            if not amp.flags["C_CONTIGUOUS"]:
                errors.append("amplitudes must be C-contiguous")
            if np.any(~np.isfinite(amp)):
                errors.append("amplitudes contains NaN or Inf values")
        else:  # torch tensor
            import torch

            if not amp.is_contiguous():
                errors.append("amplitudes must be C-contiguous")
            if not bool(torch.isfinite(amp).all()):
                errors.append("amplitudes contains NaN or Inf values")
        return errors

    def ensure_contiguous(self) -> "TraceBlock":
        """Copy with every array contiguous (base.py:102-116)."""

        def c(a):
            if a is None:
                return None
            if isinstance(a, np.ndarray):
                return np.ascontiguousarray(a)
            return a.contiguous()

        return TraceBlock(
            amplitudes=c(self.amplitudes), source_x=c(self.source_x),
            source_y=c(self.source_y), receiver_x=c(self.receiver_x),
            receiver_y=c(self.receiver_y), offset=c(self.offset),
            midpoint_x=c(self.midpoint_x), midpoint_y=c(self.midpoint_y),
            weights=c(self.weights), sample_rate_ms=self.sample_rate_ms,
            start_time_ms=self.start_time_ms)


@dataclass
class OutputTile:
    """Output tile (base.py:119-177): image (nx, ny, nt) float64 and fold
    (nx, ny, nt) int32, both accumulated by the kernel."""

    image: NDArray[np.float64]
    fold: NDArray[np.int32]
    x_axis: NDArray[np.float64]
    y_axis: NDArray[np.float64]
    t_axis_ms: NDArray[np.float64]
    x_grid: NDArray[np.float64] | None = None
    y_grid: NDArray[np.float64] | None = None

    @property
    def nx(self) -> int:
        return len(self.x_axis)

    @property
    def ny(self) -> int:
        return len(self.y_axis)

    @property
    def nt(self) -> int:
        return len(self.t_axis_ms)

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.nx, self.ny, self.nt)

    @property
    def x_range(self) -> tuple[float, float]:
        return (float(self.x_axis[0]), float(self.x_axis[-1]))

    @property
    def y_range(self) -> tuple[float, float]:
        return (float(self.y_axis[0]), float(self.y_axis[-1]))

    @property
    def dt_ms(self) -> float:
        if len(self.t_axis_ms) < 2:
            return 1.0
        return float(self.t_axis_ms[1] - self.t_axis_ms[0])

    x_coords = property(lambda self: self.x_axis)
    y_coords = property(lambda self: self.y_axis)
    t_coords_ms = property(lambda self: self.t_axis_ms)

    def reset(self) -> None:
        self.image[:] = 0.0
        self.fold[:] = 0


@dataclass
class VelocitySlice:
    """RMS velocity for one tile (base.py:180-194): (nt,) when ``is_1d``,
    otherwise (nx, ny, nt) on the tile's output grid."""

    vrms: NDArray[np.float64]
    is_1d: bool = True
    x_axis: NDArray[np.float64] | None = None
    y_axis: NDArray[np.float64] | None = None
    t_axis_ms: NDArray[np.float64] | None = None

    def get_velocity_at(self, ix: int, iy: int, it: int) -> float:
        if self.is_1d:
            return float(self.vrms[it])
        return float(self.vrms[ix, iy, it])

    def get_velocity_1d(self) -> NDArray[np.float64]:
        """Centre-pillar profile of a 3-D slice.

        The reference's Numba backends call this method on 3-D slices but
        VelocitySlice does not define it there (numba_cpu_optimized.py:570-571
        raises AttributeError); the HIP backend uses the full 3-D field
        instead and never calls this.
        """
        v = np.asarray(self.vrms)
        if self.is_1d:
            return v
        return v[v.shape[0] // 2, v.shape[1] // 2, :]


@dataclass
class KernelConfig:
    """Kernel configuration (base.py:197-265); ``None`` fields take the
    built-in defaults of the reference's settings (settings.py:94-132)."""

    max_aperture_m: float | None = None
    min_aperture_m: float | None = None
    max_dip_degrees: float | None = None
    taper_fraction: float | None = None

    aa_enabled: bool | None = None
    aa_filter_indices: NDArray[np.int32] | None = None
    aa_filter_bank: NDArray[np.float32] | None = None

    apply_spreading: bool | None = None
    apply_obliquity: bool = True

    interpolation_method: str | None = None
    output_dt_ms: float | None = None
    accumulate_fold: bool = True

    time_variant_enabled: bool = False
    time_variant_windows: list | None = None

    kernel_type: str = "straight_ray"

    curved_ray_enabled: bool = False
    curved_ray_v0: float = 1500.0
    curved_ray_k: float = 0.5

    vti_enabled: bool = False
    vti_eta_constant: float = 0.0
    vti_eta_array: NDArray[np.float64] | None = None
    vti_eta_is_1d: bool = True

    grid_dx: float = 25.0
    grid_dy: float = 12.5

    # additions of this backend
    antialias: str = "none"          # "none" | "triangle"
    aa_max_half_width: int = 16      # triangle half width clamp, samples

    def __post_init__(self) -> None:
        d = DEFAULTS
        if self.max_aperture_m is None:
            self.max_aperture_m = d.max_aperture_m
        if self.min_aperture_m is None:
            self.min_aperture_m = d.min_aperture_m
        if self.max_dip_degrees is None:
            self.max_dip_degrees = d.max_dip_degrees
        if self.taper_fraction is None:
            self.taper_fraction = d.taper_fraction
        if self.aa_enabled is None:
            self.aa_enabled = d.enable_antialiasing
        if self.apply_spreading is None:
            self.apply_spreading = d.apply_spreading_correction
        if self.interpolation_method is None:
            self.interpolation_method = d.default_interpolation
        if self.output_dt_ms is None:
            self.output_dt_ms = d.output_dt_ms

    def get_aperture_at_time(self, t_ms: float, velocity: float) -> float:
        """clip(v * t * tan(dip), min, max) as base.py:261-265 computes it
        (note: no factor 1/2, unlike the kernels' own aperture,
        numba_cpu_optimized.py:394)."""
        t_s = t_ms / 1000.0
        aperture = velocity * t_s * np.tan(np.radians(self.max_dip_degrees))
        return float(np.clip(aperture, self.min_aperture_m, self.max_aperture_m))


@dataclass
class KernelMetrics:
    """Execution statistics (base.py:268-287)."""

    n_traces_processed: int = 0
    n_samples_output: int = 0
    compute_time_s: float = 0.0
    memory_peak_mb: float = 0.0

    @property
    def traces_per_second(self) -> float:
        return self.n_traces_processed / self.compute_time_s if self.compute_time_s > 0 else 0.0

    @property
    def samples_per_second(self) -> float:
        return self.n_samples_output / self.compute_time_s if self.compute_time_s > 0 else 0.0


@runtime_checkable
class MigrationKernel(Protocol):
    """What every backend implements (base.py:290-318)."""

    @property
    def name(self) -> str: ...

    @property
    def capabilities(self) -> set[KernelCapability]: ...

    def initialize(self, config: KernelConfig) -> None: ...

    def migrate_tile(self, traces: TraceBlock, output: OutputTile,
                     velocity: VelocitySlice, config: KernelConfig) -> KernelMetrics: ...

    def synchronize(self) -> None: ...

    def cleanup(self) -> None: ...


def create_trace_block(amplitudes, source_x, source_y, receiver_x, receiver_y,
                       sample_rate_ms: float, start_time_ms: float = 0.0,
                       weights=None) -> TraceBlock:
    """TraceBlock with offset and midpoints derived in float64 (base.py:321-357)."""
    amplitudes = np.ascontiguousarray(amplitudes, dtype=np.float32)
    sx = np.asarray(source_x, dtype=np.float64)
    sy = np.asarray(source_y, dtype=np.float64)
    rx = np.asarray(receiver_x, dtype=np.float64)
    ry = np.asarray(receiver_y, dtype=np.float64)
    return TraceBlock(
        amplitudes=amplitudes, source_x=sx, source_y=sy, receiver_x=rx,
        receiver_y=ry, offset=np.sqrt((rx - sx) ** 2 + (ry - sy) ** 2),
        midpoint_x=(sx + rx) / 2.0, midpoint_y=(sy + ry) / 2.0,
        weights=None if weights is None else np.asarray(weights, dtype=np.float64),
        sample_rate_ms=sample_rate_ms, start_time_ms=start_time_ms)


def create_output_tile(x_min: float, x_max: float, dx: float, y_min: float,
                       y_max: float, dy: float, t_min_ms: float, t_max_ms: float,
                       dt_ms: float) -> OutputTile:
    """Zeroed tile on closed axis ranges (base.py:360-384)."""
    x = np.arange(x_min, x_max + dx / 2, dx, dtype=np.float64)
    y = np.arange(y_min, y_max + dy / 2, dy, dtype=np.float64)
    t = np.arange(t_min_ms, t_max_ms + dt_ms / 2, dt_ms, dtype=np.float64)
    return OutputTile(image=np.zeros((len(x), len(y), len(t)), dtype=np.float64),
                      fold=np.zeros((len(x), len(y), len(t)), dtype=np.int32),
                      x_axis=x, y_axis=y, t_axis_ms=t)
