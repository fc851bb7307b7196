"""The ``hip`` backend: MigrationKernel on MI355X through the C ABI.

Implements the reference Protocol (pstm/kernels/base.py:290-318) plus the two
optional entry points the executor looks for with ``hasattr``
(migrate_tile_curved_ray / migrate_tile_vti, pstm/pipeline/executor.py:
1710-1740).  Numerics follow the registered Numba backend
(pstm/kernels/numba_cpu_optimized.py:545-626): 1-D velocity is interpolated
onto the output time axis when its length differs (:575-578), trace weights
default to 1 (:565-567), and image and fold are ACCUMULATED into the caller's
arrays (:229,231).  Unlike the Numba backend, 3-D velocity slices are used as
full (nx, ny, nt) fields.

PyTorch is used for device memory and the stream; every computation runs in
the HIP library.  There is no CPU fallback: without a GPU or without the built
library the backend raises.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Any

import numpy as np

from pm_amd.kernels import _abi
from pm_amd.lattice import as_lattice
from pm_amd.kernels.base import (
    KernelCapability,
    KernelConfig,
    KernelMetrics,
    OutputTile,
    TraceBlock,
    VelocitySlice,
)

_KERNEL_TYPES = {"straight_ray": _abi.TT_STRAIGHT, "curved_ray": _abi.TT_CURVED,
                 "anisotropic_vti": _abi.TT_VTI}


def _torch():
    import torch

    return torch


def _dev(a: Any, dtype, device) -> Any:
    """numpy array or tensor -> contiguous device tensor of `dtype`."""
    torch = _torch()
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        t = a.to(device=device, dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)
    return t.contiguous()


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


@dataclass
class DeviceTraces:
    """A trace block resident in HBM (absolute f64 geometry, f32 samples)."""

    amplitudes: Any
    source_x: Any
    source_y: Any
    receiver_x: Any
    receiver_y: Any
    midpoint_x: Any
    midpoint_y: Any
    weights: Any
    sample_rate_ms: float
    start_time_ms: float

    @property
    def n_rows(self) -> int:
        return int(self.amplitudes.shape[0])

    @property
    def n_samples(self) -> int:
        return int(self.amplitudes.shape[1])

    @classmethod
    def from_block(cls, tb: TraceBlock, device) -> "DeviceTraces":
        torch = _torch()
        f64 = torch.float64
        return cls(_dev(tb.amplitudes, torch.float32, device),
                   _dev(tb.source_x, f64, device), _dev(tb.source_y, f64, device),
                   _dev(tb.receiver_x, f64, device), _dev(tb.receiver_y, f64, device),
                   _dev(tb.midpoint_x, f64, device), _dev(tb.midpoint_y, f64, device),
                   _dev(tb.weights, f64, device), float(tb.sample_rate_ms),
                   float(tb.start_time_ms))


class HipKernel:
    """MI355X migration kernel (registered as ComputeBackend.HIP)."""

    def __init__(self, device: int = 0):
        self._device_index = device
        self._config: KernelConfig | None = None
        self._initialized = False
        self._lib: ctypes.CDLL | None = None
        self.last_effective_contributions: int | None = None

    # -- Protocol ----------------------------------------------------------
    @property
    def name(self) -> str:
        return "hip"

    @property
    def capabilities(self) -> set[KernelCapability]:
        return {KernelCapability.ASYNC, KernelCapability.BATCH}

    def initialize(self, config: KernelConfig) -> None:
        torch = _torch()
        lib = _abi.load_library()
        if not torch.cuda.is_available():
            raise RuntimeError("hip backend: no GPU device is visible")
        torch.cuda.set_device(self._device_index)
        self._check(lib, lib.pstm_hip_init(self._device_index), "pstm_hip_init")
        self._lib = lib
        self._config = config
        self._initialized = True

    @property
    def device(self):
        return _torch().device("cuda", self._device_index)

    def migrate_tile(self, traces: TraceBlock, output: OutputTile,
                     velocity: VelocitySlice, config: KernelConfig | None = None) -> KernelMetrics:
        return self._migrate(traces, output, velocity, config, None)

    def migrate_tile_curved_ray(self, traces: TraceBlock, output: OutputTile,
                                velocity: VelocitySlice,
                                config: KernelConfig | None = None) -> KernelMetrics:
        return self._migrate(traces, output, velocity, config, _abi.TT_CURVED)

    def migrate_tile_vti(self, traces: TraceBlock, output: OutputTile,
                         velocity: VelocitySlice,
                         config: KernelConfig | None = None) -> KernelMetrics:
        return self._migrate(traces, output, velocity, config, _abi.TT_VTI)

    def migrate_tile_time_variant(self, traces: TraceBlock, output: OutputTile, velocity: VelocitySlice,
                                  config: KernelConfig | None, windows) -> KernelMetrics:
        """Migrate only the samples of the time windows
        (pm_amd.time_variant.compute_time_windows): ``output.image`` is
        replaced by the (nx, ny, n_window_samples) image — f64 NumPy, or the
        f32 device tensor for a device-resident tile — for the caller to
        resample (``pm_amd.time_variant.resample_to_uniform``), as the
        reference's executor does with its Metal kernel
        (pstm/pipeline/executor.py:1710-1726,
        pstm/kernels/metal_compiled.py:940-1230); the pillar fold is added to
        ``output.fold``.  Exact straight-ray, curved-ray and VTI traveltimes."""
        if not windows:
            raise ValueError("time-variant migration needs at least one window")
        return self._migrate(traces, output, velocity, config, None, windows=windows)

    def resample_uniform(self, tv_image, tv_times_ms, out_times_ms, out=None):
        """A (..., n_tv) f32 device image on the time-variant axis
        ``tv_times_ms`` resampled onto ``out_times_ms`` as np.interp does
        (f64 arithmetic, f32 result): a new (..., n_out) tensor, or ``out``."""
        self._require_init()
        torch = _torch()
        lib = self._lib
        if tv_image.dtype != torch.float32 or not tv_image.is_cuda or not tv_image.is_contiguous():
            raise ValueError("tv_image must be a contiguous float32 device tensor")
        xp = np.asarray(tv_times_ms, dtype=np.float64)
        x = np.asarray(out_times_ms, dtype=np.float64)
        n_tv = int(tv_image.shape[-1])
        if xp.shape != (n_tv,):
            raise ValueError(f"tv_times_ms has {xp.size} samples, the image {n_tv}")
        if n_tv > 1 and np.any(np.diff(xp) <= 0):
            raise ValueError("tv_times_ms must increase")
        shape = tuple(tv_image.shape[:-1]) + (x.size,)
        if out is None:
            out = torch.empty(shape, dtype=torch.float32, device=self.device)
        elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float32 tensor of shape {shape}")
        n_pil = int(np.prod(shape[:-1], dtype=np.int64))
        xp_d = _dev(xp, torch.float64, self.device)
        x_d = _dev(x, torch.float64, self.device)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self._device_index).cuda_stream)
        rc = lib.pstm_hip_resample_uniform(_ptr(tv_image), _ptr(xp_d), n_tv, _ptr(x_d), int(x.size), n_pil,
                                           _ptr(out), stream)
        self._check(lib, rc, "pstm_hip_resample_uniform")
        return out

    def synchronize(self) -> None:
        if self._lib is not None:
            stream = _torch().cuda.current_stream(self._device_index).cuda_stream
            self._check(self._lib, self._lib.pstm_hip_stream_sync(ctypes.c_void_p(stream)),
                        "pstm_hip_stream_sync")

    def cleanup(self) -> None:
        if self._lib is not None and self._initialized:
            self._check(self._lib, self._lib.pstm_hip_release(), "pstm_hip_release")
        self._initialized = False
        self._config = None

    # -- device-resident API -------------------------------------------------
    def upload_traces(self, traces: TraceBlock) -> DeviceTraces:
        self._require_init()
        return DeviceTraces.from_block(traces, self.device)

    def migrate_device(self, dt: DeviceTraces, image, fold2d, x_axis, y_axis, t_axis_ms,
                       vrms, config: KernelConfig, *, traveltime: int | None = None,
                       trace_index=None, eta=None, count_effective: bool = False,
                       group_start: np.ndarray | None = None, surrogate=None,
                       node_spacing=16) -> int | None:
        """Accumulate one tile into device tensors ``image`` (nx, ny, nt) f32
        and ``fold2d`` (nx, ny) i32.  All array arguments are device tensors
        (vrms (nt,) or (nx, ny, nt) f64).  Asynchronous on the current stream
        unless ``count_effective`` is set, in which case the effective
        contribution count E is returned.  ``traveltime=TT_VTI_LEARNED`` needs
        ``surrogate`` (a trained pm_amd.surrogate.TraveltimeSurrogate) and
        evaluates it on a t0 lattice inside the gather: ``node_spacing`` is a
        pm_amd.lattice.Lattice or an int (uniform spacing in output samples)."""
        self._require_init()
        torch = _torch()
        lib = self._lib
        counter = torch.zeros(1, dtype=torch.int64, device=self.device) if count_effective else None
        a, keep = self._args(dt, image, fold2d, x_axis, y_axis, t_axis_ms, vrms, config, traveltime,
                             trace_index, eta, counter, surrogate, node_spacing)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self._device_index).cuda_stream)
        if group_start is None:
            rc = lib.pstm_hip_migrate_tile(ctypes.byref(a), stream)
            self._check(lib, rc, "pstm_hip_migrate_tile")
        else:
            gs = np.ascontiguousarray(group_start, dtype=np.int32)
            rc = lib.pstm_hip_migrate_bins(ctypes.byref(a), gs.ctypes.data_as(ctypes.c_void_p),
                                           int(gs.shape[0] - 1), stream)
            self._check(lib, rc, "pstm_hip_migrate_bins")
        del keep
        if counter is not None:
            self.synchronize()
            return int(counter.item())
        return None

    def create_plan(self, dt: DeviceTraces, tiles: list[dict], t_axis_ms, config: KernelConfig, *,
                    traveltime: int | None = None, count_effective: bool = False, surrogate=None,
                    node_spacing=16) -> "HipPlan":
        """A device-resident work queue over many output tiles: one persistent,
        hipGraph-captured gather per run (``HipPlan.run``).  Each tile is a
        dict with image, fold (device tensors, accumulated), x_axis, y_axis,
        vrms, trace_index and optionally eta; ``dt`` is the whole trace block
        (every tile's trace_index indexes its rows)."""
        self._require_init()
        torch = _torch()
        lib = self._lib
        if not tiles:
            raise ValueError("a plan needs at least one tile")
        counter = torch.zeros(1, dtype=torch.int64, device=self.device) if count_effective else None
        t0 = tiles[0]
        a, keep = self._args(dt, t0["image"], t0["fold"], t0["x_axis"], t0["y_axis"], t_axis_ms, t0["vrms"],
                             config, traveltime, t0["trace_index"], t0.get("eta"), counter, surrogate,
                             node_spacing)
        a.n_traces = dt.n_rows
        descs = (_abi.PstmTileDesc * len(tiles))()
        nt = int(t_axis_ms.shape[0])
        for d, t in zip(descs, tiles):
            nx, ny = int(t["image"].shape[0]), int(t["image"].shape[1])
            if tuple(t["image"].shape) != (nx, ny, nt) or t["image"].dtype != torch.float32:
                raise ValueError("tile image must be (nx, ny, nt) float32")
            v = t["vrms"]
            if (v.dim() == 3) != (a.vel_mode == _abi.VEL_3D):
                raise ValueError("every tile of a plan needs the same velocity layout")
            if v.dim() == 3 and tuple(v.shape) != (nx, ny, nt):
                raise ValueError(f"3-D velocity shape {tuple(v.shape)} != tile {(nx, ny, nt)}")
            eta = t.get("eta")
            if (eta is None) != (a.eta_mode == _abi.ETA_CONST) or (
                    eta is not None and (eta.dim() == 1) != (a.eta_mode == _abi.ETA_1D)):
                raise ValueError("every tile of a plan needs the same eta layout")
            d.x_axis, d.y_axis, d.vrms = _ptr(t["x_axis"]), _ptr(t["y_axis"]), _ptr(v)
            d.eta = _ptr(eta) if eta is not None else None
            d.image, d.fold = _ptr(t["image"]), _ptr(t["fold"])
            d.trace_index = _ptr(t["trace_index"])
            d.n_traces = int(t["trace_index"].shape[0])
            d.nx, d.ny = nx, ny
            keep.extend([t["x_axis"], t["y_axis"], v, eta, t["image"], t["fold"], t["trace_index"]])
        handle = ctypes.c_void_p()
        rc = lib.pstm_hip_plan_create(ctypes.byref(a), descs, len(tiles), ctypes.byref(handle))
        self._check(lib, rc, "pstm_hip_plan_create")
        return HipPlan(self, handle, keep, counter)

    # -- output step ---------------------------------------------------------
    def bin_headers(self, dt: DeviceTraces, rows, x_min: float, y_min: float, dx: float, dy: float,
                    x_start: int, y_start: int, tile_nx: int, tile_ny: int, sums) -> None:
        """Add one tile's per-bin trace-header sums (count, offset, azimuth
        sine / cosine) into ``sums`` (a pm_amd.pipeline.output.HeaderSums on
        this device), as the reference's executor does after each tile
        (pstm/pipeline/executor.py:1783-1806).  ``rows``: int32 device rows of
        ``dt`` (the tile's trace query), or None for every row.  Asynchronous
        on the current stream."""
        self._require_init()
        torch = _torch()
        lib = self._lib
        if rows is not None and (rows.dtype != torch.int32 or not rows.is_cuda):
            raise ValueError("rows must be an int32 device tensor")
        n = dt.n_rows if rows is None else int(rows.shape[0])
        gnx, gny = sums.shape
        stream = ctypes.c_void_p(torch.cuda.current_stream(self._device_index).cuda_stream)
        rc = lib.pstm_hip_bin_headers(_ptr(dt.source_x), _ptr(dt.source_y), _ptr(dt.receiver_x), _ptr(dt.receiver_y),
                                      _ptr(rows) if rows is not None else None, n, float(x_min), float(y_min),
                                      float(dx), float(dy), int(x_start), int(y_start), int(tile_nx), int(tile_ny),
                                      gnx, gny, _ptr(sums.trace_count), _ptr(sums.offset_sum),
                                      _ptr(sums.azimuth_sin_sum), _ptr(sums.azimuth_cos_sum), stream)
        self._check(lib, rc, "pstm_hip_bin_headers")

    def select_rows(self, dt: DeviceTraces, rows, gather_bin):
        """The rows (int32 device tensor, in their order) of ``rows`` (None:
        every row of ``dt``) that pass ``gather_bin``'s offset or
        offset-vector-tile filter (executor.py:1537-1554; closed bounds, a
        bound of None is not checked)."""
        self._require_init()
        torch = _torch()
        lib = self._lib
        if rows is None:
            rows = torch.arange(dt.n_rows, dtype=torch.int32, device=self.device)
        if rows.dtype != torch.int32 or not rows.is_cuda:
            raise ValueError("rows must be an int32 device tensor")
        n = int(rows.shape[0])
        keep = torch.empty(n, dtype=torch.uint8, device=self.device)
        checks, bounds = 0, []
        for bit, v in ((_abi.SEL_OFFSET_MIN, gather_bin.offset_min), (_abi.SEL_OFFSET_MAX, gather_bin.offset_max),
                       (_abi.SEL_OVT_X_MIN, gather_bin.ovt_x_min), (_abi.SEL_OVT_X_MAX, gather_bin.ovt_x_max),
                       (_abi.SEL_OVT_Y_MIN, gather_bin.ovt_y_min), (_abi.SEL_OVT_Y_MAX, gather_bin.ovt_y_max)):
            if v is not None:
                checks |= bit
            bounds.append(0.0 if v is None else float(v))
        stream = ctypes.c_void_p(torch.cuda.current_stream(self._device_index).cuda_stream)
        rc = lib.pstm_hip_select_rows(_ptr(dt.source_x), _ptr(dt.source_y), _ptr(dt.receiver_x), _ptr(dt.receiver_y),
                                      _ptr(rows), n, checks, *bounds, _ptr(keep), stream)
        self._check(lib, rc, "pstm_hip_select_rows")
        return rows[keep.bool()]

    def fold_normalize(self, image, fold2d, out=None):
        """image (nx, ny, nt) f32 divided by the pillar fold (nx, ny) i32 where
        it is positive, 0 elsewhere (executor.py:2155-2160), into ``out``
        (a new tensor by default; may be ``image``).  Returns ``out``."""
        self._require_init()
        torch = _torch()
        lib = self._lib
        if image.dtype != torch.float32 or image.dim() != 3 or not image.is_contiguous():
            raise ValueError("image must be a contiguous (nx, ny, nt) float32 device tensor")
        if fold2d.dtype != torch.int32 or tuple(fold2d.shape) != tuple(image.shape[:2]) or not fold2d.is_contiguous():
            raise ValueError("fold must be a contiguous (nx, ny) int32 device tensor matching the image")
        if out is None:
            out = torch.empty_like(image)
        elif out.shape != image.shape or out.dtype != torch.float32 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float32 tensor shaped like image")
        nx, ny, nt = (int(s) for s in image.shape)
        stream = ctypes.c_void_p(torch.cuda.current_stream(self._device_index).cuda_stream)
        rc = lib.pstm_hip_fold_normalize(_ptr(image), _ptr(fold2d), nx * ny, nt, _ptr(out), stream)
        self._check(lib, rc, "pstm_hip_fold_normalize")
        return out

    def _args(self, dt, image, fold2d, x_axis, y_axis, t_axis_ms, vrms, config, traveltime, trace_index, eta,
              counter, surrogate, node_spacing):
        nx, ny, nt = (int(s) for s in image.shape)
        tt = self._traveltime(config) if traveltime is None else traveltime
        a = _abi.PstmMigrateArgs()
        keep: list[Any] = [dt, image, fold2d, x_axis, y_axis, t_axis_ms, vrms, counter]
        a.amplitudes = _ptr(dt.amplitudes)
        a.source_x, a.source_y = _ptr(dt.source_x), _ptr(dt.source_y)
        a.receiver_x, a.receiver_y = _ptr(dt.receiver_x), _ptr(dt.receiver_y)
        a.midpoint_x, a.midpoint_y = _ptr(dt.midpoint_x), _ptr(dt.midpoint_y)
        a.trace_weights = _ptr(dt.weights)
        if trace_index is not None:
            keep.append(trace_index)
            a.trace_index = _ptr(trace_index)
            a.n_traces = int(trace_index.shape[0])
        else:
            a.n_traces = dt.n_rows
        a.x_axis, a.y_axis, a.t_axis_ms = _ptr(x_axis), _ptr(y_axis), _ptr(t_axis_ms)
        a.vrms = _ptr(vrms)
        a.vel_mode = _abi.VEL_1D if vrms.dim() == 1 else _abi.VEL_3D
        if a.vel_mode == _abi.VEL_3D and tuple(vrms.shape) != (nx, ny, nt):
            raise ValueError(f"3-D velocity shape {tuple(vrms.shape)} != tile {(nx, ny, nt)}")
        if a.vel_mode == _abi.VEL_1D and int(vrms.shape[0]) != nt:
            raise ValueError(f"1-D velocity length {int(vrms.shape[0])} != nt {nt}")
        if eta is not None:
            keep.append(eta)
            a.eta = _ptr(eta)
            a.eta_mode = _abi.ETA_1D if eta.dim() == 1 else _abi.ETA_3D
        else:
            a.eta_mode = _abi.ETA_CONST
        a.image, a.fold = _ptr(image), _ptr(fold2d)
        if counter is not None:
            a.n_effective = _ptr(counter)
        a.n_samples = dt.n_samples
        a.nx, a.ny, a.nt = nx, ny, nt
        a.traveltime = tt
        aa = (config.antialias or "none").lower()
        if aa not in ("none", "triangle"):
            raise ValueError(f"unknown antialias method {config.antialias!r}")
        a.aa_mode = _abi.AA_TRIANGLE if aa == "triangle" else _abi.AA_NONE
        a.aa_max_half_width = int(config.aa_max_half_width)
        # KernelConfig.interpolation_method, as the reference's NumbaKernel reads
        # it (pstm/kernels/numba_cpu.py:632); straight ray without AA only
        a.interpolation = _abi.interp_code(config.interpolation_method)
        a.apply_spreading = int(bool(config.apply_spreading))
        a.apply_obliquity = int(bool(config.apply_obliquity))
        a.dt_in_ms, a.t_start_in_ms = dt.sample_rate_ms, dt.start_time_ms
        a.max_aperture_m = float(config.max_aperture_m)
        a.min_aperture_m = float(config.min_aperture_m)
        a.max_dip_deg = float(config.max_dip_degrees)
        a.taper_fraction = float(config.taper_fraction)
        a.grid_dx, a.grid_dy = float(config.grid_dx), float(config.grid_dy)
        a.curved_v0, a.curved_k = float(config.curved_ray_v0), float(config.curved_ray_k)
        a.eta_constant = float(config.vti_eta_constant)
        if tt in (_abi.TT_VTI_LEARNED, _abi.TT_CURVED_LEARNED):
            if surrogate is None:
                raise ValueError("the learned traveltime needs a trained surrogate")
            n = surrogate.norm
            keep.append(surrogate.params)
            a.mlp_params = _ptr(surrogate.params)
            lat = as_lattice(node_spacing)
            a.mlp_node_spacing, a.mlp_node_growth, a.mlp_node_max = lat.spacing, lat.growth, lat.max_spacing
            a.mlp_tau_max, a.mlp_v_lo, a.mlp_v_hi = float(n.tau_max), float(n.v_lo), float(n.v_hi)
            a.mlp_eta_max, a.mlp_out_scale = float(n.eta_max), float(n.out_scale)
            a.mlp_weight_terms = int(getattr(surrogate, "weight_terms", 1))
        return a, keep

    # -- internals -----------------------------------------------------------
    def _require_init(self) -> None:
        if not self._initialized or self._config is None:
            raise RuntimeError("Kernel not initialized")

    @staticmethod
    def _check(lib, rc: int, what: str) -> None:
        if rc != 0:
            msg = _abi.last_error(lib)
            if rc == 1:
                raise ValueError(f"{what}: {msg}")
            raise RuntimeError(f"{what} failed ({rc}): {msg}")

    @staticmethod
    def _traveltime(config: KernelConfig) -> int:
        if config.vti_enabled:
            return _abi.TT_VTI
        if config.curved_ray_enabled:
            return _abi.TT_CURVED
        try:
            return _KERNEL_TYPES[config.kernel_type]
        except KeyError:
            raise ValueError(f"unknown kernel_type {config.kernel_type!r}") from None

    def _migrate(self, traces: TraceBlock, output: OutputTile, velocity: VelocitySlice,
                 config: KernelConfig | None, traveltime: int | None, windows=None) -> KernelMetrics:
        self._require_init()
        torch = _torch()
        cfg = config if config is not None else self._config
        start = time.perf_counter()
        dev = self.device
        nx, ny, nt = output.nx, output.ny, output.nt
        t_axis = np.asarray(output.t_axis_ms, dtype=np.float64)
        t_index = None
        if windows is not None:
            from pm_amd.time_variant import sample_times, uniform_index

            t_tv = sample_times(windows)
            t_index = uniform_index(t_tv, t_axis)
        # velocity onto the output grid (numba_cpu_optimized.py:569-578)
        if velocity.is_1d:
            v = np.asarray(velocity.vrms, dtype=np.float64)
            if v.ndim != 1:
                raise ValueError("1-D VelocitySlice must hold a (nt,) array")
            if len(v) != nt:
                vt = velocity.t_axis_ms if velocity.t_axis_ms is not None else t_axis
                v = np.interp(t_axis, np.asarray(vt, dtype=np.float64), v)
        else:
            v = velocity.vrms
            if tuple(v.shape) != (nx, ny, nt):
                raise ValueError(
                    f"3-D VelocitySlice shape {tuple(v.shape)} does not match tile {(nx, ny, nt)}")
        tt = self._traveltime(cfg) if traveltime is None else traveltime
        eta = None
        if tt == _abi.TT_VTI and cfg.vti_eta_array is not None:
            eta = np.asarray(cfg.vti_eta_array, dtype=np.float64)
        if t_index is not None:
            # the window samples, with the velocity (and eta) of the uniform
            # sample each one falls on
            t_axis = t_tv
            v = np.ascontiguousarray(np.asarray(v)[..., t_index])
            if eta is not None and eta.ndim >= 1 and eta.shape[-1] == nt:
                eta = np.ascontiguousarray(eta[..., t_index])
            nt = len(t_tv)
        if eta is not None:
            eta = _dev(eta, torch.float64, dev)

        dt = DeviceTraces.from_block(traces, dev)
        f64 = torch.float64
        on_device = isinstance(output.image, torch.Tensor) and output.image.is_cuda
        if on_device and t_index is None:
            image = output.image
            if image.dtype != torch.float32:
                raise ValueError("device-resident OutputTile.image must be float32")
        else:
            image = torch.zeros((nx, ny, nt), dtype=torch.float32, device=dev)
        fold2d = torch.zeros((nx, ny), dtype=torch.int32, device=dev)
        self.migrate_device(dt, image, fold2d, _dev(output.x_axis, f64, dev),
                            _dev(output.y_axis, f64, dev), _dev(t_axis, f64, dev),
                            _dev(v, f64, dev), cfg, traveltime=tt, eta=eta)
        if on_device:
            if t_index is not None:
                output.image = image
            f = output.fold
            f += fold2d.view(nx, ny, 1) if f.dim() == 3 else fold2d
        else:
            img_h = image.cpu().numpy()
            fold_h = fold2d.cpu().numpy()
            if t_index is not None:
                output.image = img_h.astype(np.float64)
            else:
                output.image += img_h
            if output.fold.ndim == 3:
                output.fold += fold_h[:, :, None]
            else:
                output.fold += fold_h
        return KernelMetrics(n_traces_processed=traces.n_traces,
                             n_samples_output=nx * ny * nt,
                             compute_time_s=time.perf_counter() - start)


class HipPlan:
    """A plan built by HipKernel.create_plan: run() migrates every tile once
    (image and fold +=), asynchronously on the current stream."""

    def __init__(self, kernel: HipKernel, handle, keep: list, counter):
        self.kernel = kernel
        self._h = handle
        self._keep = keep          # device tensors the plan's pointers refer to
        self.counter = counter     # E accumulates here when the plan counts

    def run(self, use_graph: bool = True) -> None:
        torch = _torch()
        if self._h is None:
            raise RuntimeError("plan destroyed")
        lib = self.kernel._lib
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.kernel._device_index).cuda_stream)
        HipKernel._check(lib, lib.pstm_hip_plan_run(self._h, 1 if use_graph else 0, stream), "pstm_hip_plan_run")

    def info(self) -> dict:
        lib = self.kernel._lib
        n = ctypes.c_longlong()
        g, mw, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        HipKernel._check(lib, lib.pstm_hip_plan_info(self._h, ctypes.byref(n), ctypes.byref(g), ctypes.byref(mw),
                                                     ctypes.byref(b)), "pstm_hip_plan_info")
        return {"items": int(n.value), "grid": g.value, "mw": mw.value, "batch": b.value}

    def destroy(self) -> None:
        if self._h is not None:
            self.kernel._lib.pstm_hip_plan_destroy(self._h)
            self._h = None
            self._keep = []

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def create_hip_kernel(device: int = 0) -> HipKernel:
    return HipKernel(device=device)
