"""Time-variant output sampling: migrate at a coarser output dt where the
data hold no high frequencies, then interpolate back to the uniform axis.

The reference plans windows on the host from a (time, f_max) table
(pstm/algorithm/time_variant.py:19-287), migrates only the window samples
(Metal ``pstm_migrate_time_variant``, pstm/metal/shaders/
pstm_migration.metal:544-701, called through
pstm/kernels/metal_compiled.py:940-1230) and resamples each pillar with
``np.interp`` (time_variant.py:354-397, called at
pstm/pipeline/executor.py:1710-1726).  Here:

  * the window plan is the reference's (same steps, so the same windows):
    f_max(t) by monotone piecewise-cubic Hermite interpolation of the table
    (SciPy's PchipInterpolator, the reference's own dependency), held at the
    table's end values outside it; the downsample factor is the Nyquist
    interval over the base dt, rounded to the nearer power of two (ties
    down) and clamped; a window keeps its factor for as long as the factor
    at the next sample would be the same;
  * the window samples are migrated by the same gather as the uniform axis
    (every kernel takes an arbitrary output time axis), with the velocity
    (and eta) of the uniform sample each window sample falls on — so a
    time-variant image equals the uniform image at the window samples;
  * the resampling runs on the device (``pstm_hip_resample_uniform``) with
    the arithmetic of ``np.interp`` in f64, so it is bit-exact with it.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

#: the reference's TimeVariantConfig default (pstm/config/models.py:413-420)
CONFIG_DEFAULT_TABLE = ((0.0, 80.0), (1000.0, 50.0), (2500.0, 30.0), (5000.0, 20.0))


@dataclass
class FrequencyTimeTable:
    """(time ms, maximum frequency Hz) pairs; times strictly increasing,
    frequencies positive, at least two entries (time_variant.py:19-86)."""

    times_ms: list = field(default_factory=list)
    frequencies_hz: list = field(default_factory=list)

    def __post_init__(self):
        self.times_ms = [float(t) for t in self.times_ms]
        self.frequencies_hz = [float(f) for f in self.frequencies_hz]
        if len(self.times_ms) != len(self.frequencies_hz):
            raise ValueError("times and frequencies must have the same length")
        if len(self.times_ms) < 2:
            raise ValueError("a frequency table needs at least 2 entries")
        if any(b <= a for a, b in zip(self.times_ms, self.times_ms[1:])):
            raise ValueError(f"times must increase strictly: {self.times_ms}")
        if any(f <= 0 for f in self.frequencies_hz):
            raise ValueError(f"frequencies must be positive: {self.frequencies_hz}")

    @classmethod
    def from_pairs(cls, pairs) -> "FrequencyTimeTable":
        """From (time, frequency) pairs in any order."""
        pairs = sorted(((float(t), float(f)) for t, f in pairs), key=lambda p: p[0])
        if not pairs:
            raise ValueError("no (time, frequency) pairs")
        return cls([t for t, _ in pairs], [f for _, f in pairs])

    @classmethod
    def default(cls) -> "FrequencyTimeTable":
        """The reference's built-in table (time_variant.py:64-70)."""
        return cls([0.0, 500.0, 1500.0, 3000.0, 5000.0], [80.0, 60.0, 40.0, 25.0, 15.0])

    def fmax(self, t_ms: float) -> float:
        """Maximum frequency at t: the end values outside the table, the
        monotone cubic (PCHIP) inside."""
        if t_ms <= self.times_ms[0]:
            return self.frequencies_hz[0]
        if t_ms >= self.times_ms[-1]:
            return self.frequencies_hz[-1]
        if not hasattr(self, "_pchip"):
            from scipy.interpolate import PchipInterpolator

            self._pchip = PchipInterpolator(self.times_ms, self.frequencies_hz)
        return float(self._pchip(t_ms))


@dataclass(frozen=True)
class TimeWindow:
    """Output samples t_start_ms + i dt_effective_ms, i < n_samples, stored at
    [sample_start, sample_end) of the time-variant axis (time_variant.py:88-121)."""

    t_start_ms: float
    t_end_ms: float
    dt_effective_ms: float
    downsample_factor: int
    sample_start: int
    sample_end: int
    f_max_hz: float

    @property
    def n_samples(self) -> int:
        return self.sample_end - self.sample_start

    @property
    def duration_ms(self) -> float:
        return self.t_end_ms - self.t_start_ms


def nearest_power_of_2(n: int) -> int:
    """The power of two nearer to n (the lower one on a tie), at least 1."""
    if n <= 1:
        return 1
    lo = 1 << (int(n).bit_length() - 1)
    return lo * 2 if (n - lo) > (lo * 2 - n) else lo


def downsample_factor(f_max_hz: float, base_dt_ms: float, min_factor: int = 1, max_factor: int = 8) -> int:
    """Base samples per Nyquist interval 1000 / (2 f_max) ms, truncated,
    rounded to a power of two and clamped to [min_factor, max_factor]."""
    n = int(1000.0 / (2.0 * f_max_hz) / base_dt_ms)
    return max(min_factor, min(max_factor, nearest_power_of_2(n)))


def compute_time_windows(t_min_ms: float, t_max_ms: float, base_dt_ms: float, table: FrequencyTimeTable,
                         min_factor: int = 1, max_factor: int = 8) -> list[TimeWindow]:
    """Windows covering [t_min_ms, t_max_ms) in order (time_variant.py:201-286).

    A window starts at the current time t with the factor of f_max(t) and
    grows by its own dt while the factor at the next sample is unchanged;
    it holds at least one sample, and its end is snapped to a whole number
    of its samples (capped at t_max_ms)."""
    windows: list[TimeWindow] = []
    t, first = float(t_min_ms), 0
    while t < t_max_ms:
        f = table.fmax(t)
        fac = downsample_factor(f, base_dt_ms, min_factor, max_factor)
        step = base_dt_ms * fac
        end = t
        while end < t_max_ms and downsample_factor(table.fmax(end + step), base_dt_ms, min_factor,
                                                   max_factor) == fac:
            end += step
        if end <= t:
            end = min(t + step, t_max_ms)
        end = min(end, t_max_ms)
        n = max(1, int((end - t) / step))
        end = t + n * step
        if end > t_max_ms:
            end = t_max_ms
            n = max(1, int((end - t) / step))
        windows.append(TimeWindow(t, end, step, fac, first, first + n, f))
        first += n
        t = end
    return windows


def sample_times(windows: list[TimeWindow]) -> np.ndarray:
    """The time-variant output axis (ms), window by window."""
    if not windows:
        return np.zeros(0)
    return np.concatenate([w.t_start_ms + w.dt_effective_ms * np.arange(w.n_samples) for w in windows])


def estimate_speedup(t_min_ms: float, t_max_ms: float, base_dt_ms: float, table: FrequencyTimeTable) -> float:
    """Uniform samples over time-variant samples (time_variant.py:289-317)."""
    uniform = int((t_max_ms - t_min_ms) / base_dt_ms)
    tv = sum(w.n_samples for w in compute_time_windows(t_min_ms, t_max_ms, base_dt_ms, table))
    return uniform / tv if tv else 1.0


def uniform_index(tv_times_ms: np.ndarray, t_axis_ms: np.ndarray, tol: float = 1e-6) -> np.ndarray:
    """Index of the uniform output sample each time-variant sample falls on;
    raises ValueError if one falls between uniform samples (windows planned
    with another base dt or start)."""
    t_axis_ms = np.asarray(t_axis_ms, dtype=np.float64)
    if t_axis_ms.size < 2:
        raise ValueError("the uniform axis needs at least two samples")
    dt = t_axis_ms[1] - t_axis_ms[0]
    k = np.rint((np.asarray(tv_times_ms, dtype=np.float64) - t_axis_ms[0]) / dt).astype(np.int64)
    if k.size and (k.min() < 0 or k.max() >= t_axis_ms.size):
        raise ValueError("time-variant samples outside the uniform axis")
    if k.size and np.max(np.abs(t_axis_ms[k] - tv_times_ms)) > tol * abs(dt):
        raise ValueError("time-variant samples are not on the uniform axis")
    return k


def uniform_times(base_dt_ms: float, nt: int) -> np.ndarray:
    """The axis the reference resamples onto: k * base_dt_ms, k < nt
    (time_variant.py:386), which starts at 0 whatever the grid's t_min."""
    return np.arange(nt) * float(base_dt_ms)


def resample_to_uniform(tv_image, windows: list[TimeWindow], base_dt_ms: float, output_nt: int, kernel):
    """The time-variant image (nx, ny, n_window_samples) on the uniform axis
    k * base_dt_ms, k < output_nt, as np.interp per pillar
    (time_variant.py:354-397), on the device through ``kernel`` (an
    initialised HipKernel).  A device tensor gives a new f32 device tensor;
    a NumPy array gives an f64 NumPy array of the f32 results."""
    import torch

    xp = sample_times(windows)
    x = uniform_times(base_dt_ms, output_nt)
    if isinstance(tv_image, torch.Tensor) and tv_image.is_cuda:
        return kernel.resample_uniform(tv_image.to(torch.float32).contiguous(), xp, x)
    dev = torch.as_tensor(np.ascontiguousarray(tv_image, dtype=np.float32), device=kernel.device)
    return kernel.resample_uniform(dev, xp, x).cpu().numpy().astype(np.float64)
