"""TEST INFRASTRUCTURE: ctypes wrapper of the C++ FP64 oracle (pstm_oracle.cpp).

The oracle restates _migrate_tile_kernel_optimized and _precompute_time_terms
(pstm/kernels/numba_cpu_optimized.py:93-233,358-401) and the build's own
extensions (3-D velocity, curved ray, VTI, triangle anti-aliasing); see the
header of pstm_oracle.cpp for the reference lines each part follows.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from pm_amd.kernels._abi import (AA_NONE, AA_TRIANGLE, ETA_1D, ETA_3D, ETA_CONST,
                                 TT_STRAIGHT, VEL_1D, VEL_3D, PstmMigrateArgs)

HERE = Path(__file__).resolve().parent
BUILD = HERE / "build"
_LIBS: dict[str, ctypes.CDLL] = {}


def build() -> None:
    """Compile liboracle.so (strict IEEE) and liboracle_fast.so (-ffast-math)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _load(fast: bool = False) -> ctypes.CDLL:
    key = "fast" if fast else "strict"
    if key in _LIBS:
        return _LIBS[key]
    path = BUILD / ("liboracle_fast.so" if fast else "liboracle.so")
    if not path.exists():
        build()
    lib = ctypes.CDLL(str(path))
    lib.oracle_migrate_tile.restype = ctypes.c_int
    lib.oracle_migrate_tile.argtypes = [ctypes.POINTER(PstmMigrateArgs), ctypes.c_void_p,
                                        ctypes.c_int]
    lib.oracle_args_size.restype = ctypes.c_size_t
    for fn in ("oracle_curved_dsr", "oracle_vti_dsr"):
        getattr(lib, fn).restype = ctypes.c_double
        getattr(lib, fn).argtypes = [ctypes.c_double] * 5
    lib.oracle_vti_leg.restype = ctypes.c_double
    lib.oracle_vti_leg.argtypes = [ctypes.c_double] * 4
    lib.oracle_interp_sample.restype = ctypes.c_double
    lib.oracle_interp_sample.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.oracle_fold_tile.restype = ctypes.c_int
    lib.oracle_fold_tile.argtypes = [ctypes.POINTER(PstmMigrateArgs), ctypes.c_int]
    lib.oracle_triangle_at.restype = ctypes.c_double
    lib.oracle_triangle_at.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.oracle_integrate_traces.restype = None
    lib.oracle_integrate_traces.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p]
    if lib.oracle_args_size() != ctypes.sizeof(PstmMigrateArgs):
        raise RuntimeError("oracle/ABI struct layout mismatch")
    _LIBS[key] = lib
    return lib


@dataclass
class OracleResult:
    image: np.ndarray           # (nx, ny, nt) f64, accumulated
    fold2d: np.ndarray          # (nx, ny) i32 pillar fold
    n_effective: int            # E: contributions past aperture + bounds

    @property
    def fold3d(self) -> np.ndarray:
        return np.repeat(self.fold2d[:, :, None], self.image.shape[2], axis=2)


@dataclass
class MigrateInputs:
    """Host arrays of one tile migration, in the reference's units."""

    amplitudes: np.ndarray
    source_x: np.ndarray
    source_y: np.ndarray
    receiver_x: np.ndarray
    receiver_y: np.ndarray
    midpoint_x: np.ndarray
    midpoint_y: np.ndarray
    x_axis: np.ndarray
    y_axis: np.ndarray
    t_axis_ms: np.ndarray
    vrms: np.ndarray                 # (nt,) or (nx, ny, nt)
    dt_in_ms: float = 2.0
    t_start_in_ms: float = 0.0
    trace_weights: np.ndarray | None = None
    trace_index: np.ndarray | None = None
    eta: np.ndarray | None = None    # None, (nt,) or (nx, ny, nt)
    max_aperture_m: float = 5000.0
    min_aperture_m: float = 500.0
    max_dip_deg: float = 45.0
    taper_fraction: float = 0.1
    apply_spreading: bool = True
    apply_obliquity: bool = True
    traveltime: int = TT_STRAIGHT
    aa_mode: int = AA_NONE
    aa_max_half_width: int = 16
    grid_dx: float = 25.0
    grid_dy: float = 12.5
    curved_v0: float = 1500.0
    curved_k: float = 0.5
    eta_constant: float = 0.0
    interpolation: int = 0           # PSTM_INTERP_* (0 = linear)
    _keep: list = field(default_factory=list, repr=False)


def _p(keep: list, a: np.ndarray | None, dtype) -> int | None:
    if a is None:
        return None
    arr = np.ascontiguousarray(a, dtype=dtype)
    keep.append(arr)
    return arr.ctypes.data


def migrate(inp: MigrateInputs, image_init: np.ndarray | None = None,
            fold_init: np.ndarray | None = None, threads: int | None = None,
            fast: bool = False) -> OracleResult:
    lib = _load(fast)
    nx, ny, nt = len(inp.x_axis), len(inp.y_axis), len(inp.t_axis_ms)
    keep: list = []
    a = PstmMigrateArgs()
    a.amplitudes = _p(keep, inp.amplitudes, np.float32)
    a.source_x = _p(keep, inp.source_x, np.float64)
    a.source_y = _p(keep, inp.source_y, np.float64)
    a.receiver_x = _p(keep, inp.receiver_x, np.float64)
    a.receiver_y = _p(keep, inp.receiver_y, np.float64)
    a.midpoint_x = _p(keep, inp.midpoint_x, np.float64)
    a.midpoint_y = _p(keep, inp.midpoint_y, np.float64)
    a.trace_weights = _p(keep, inp.trace_weights, np.float64)
    a.trace_index = _p(keep, inp.trace_index, np.int32)
    a.x_axis = _p(keep, inp.x_axis, np.float64)
    a.y_axis = _p(keep, inp.y_axis, np.float64)
    a.t_axis_ms = _p(keep, inp.t_axis_ms, np.float64)
    v = np.asarray(inp.vrms, dtype=np.float64)
    a.vrms = _p(keep, v, np.float64)
    a.vel_mode = VEL_1D if v.ndim == 1 else VEL_3D
    if inp.eta is None:
        a.eta_mode = ETA_CONST
    else:
        e = np.asarray(inp.eta, dtype=np.float64)
        a.eta = _p(keep, e, np.float64)
        a.eta_mode = ETA_1D if e.ndim == 1 else ETA_3D
    fold = (np.zeros((nx, ny), np.int32) if fold_init is None
            else np.array(fold_init, dtype=np.int32, copy=True))
    a.fold = fold.ctypes.data
    counter = np.zeros(1, dtype=np.uint64)
    a.n_effective = counter.ctypes.data
    a.n_traces = (len(inp.trace_index) if inp.trace_index is not None
                  else int(inp.amplitudes.shape[0]))
    a.n_samples = int(inp.amplitudes.shape[1])
    a.nx, a.ny, a.nt = nx, ny, nt
    a.traveltime = inp.traveltime
    a.aa_mode = inp.aa_mode
    a.aa_max_half_width = inp.aa_max_half_width
    a.apply_spreading = int(inp.apply_spreading)
    a.apply_obliquity = int(inp.apply_obliquity)
    a.dt_in_ms, a.t_start_in_ms = inp.dt_in_ms, inp.t_start_in_ms
    a.max_aperture_m, a.min_aperture_m = inp.max_aperture_m, inp.min_aperture_m
    a.max_dip_deg, a.taper_fraction = inp.max_dip_deg, inp.taper_fraction
    a.grid_dx, a.grid_dy = inp.grid_dx, inp.grid_dy
    a.curved_v0, a.curved_k, a.eta_constant = inp.curved_v0, inp.curved_k, inp.eta_constant
    a.interpolation = int(inp.interpolation)
    image = (np.zeros((nx, ny, nt)) if image_init is None
             else np.array(image_init, dtype=np.float64, copy=True))
    nthr = threads if threads is not None else (os.cpu_count() or 1)
    rc = lib.oracle_migrate_tile(ctypes.byref(a), image.ctypes.data, int(nthr))
    if rc != 0:
        raise ValueError(f"oracle_migrate_tile failed ({rc})")
    return OracleResult(image=image, fold2d=fold, n_effective=int(counter[0]))


def fold(x_axis, y_axis, midpoint_x, midpoint_y, max_aperture_m: float, trace_index=None,
         threads: int | None = None) -> np.ndarray:
    """(nx, ny) i32 pillar fold of the given midpoints, the oracle's f64 test
    (strict IEEE build: the decision at the aperture edge must be the
    oracle's)."""
    lib = _load(False)
    keep: list = []
    a = PstmMigrateArgs()
    a.x_axis = _p(keep, x_axis, np.float64)
    a.y_axis = _p(keep, y_axis, np.float64)
    a.midpoint_x = _p(keep, midpoint_x, np.float64)
    a.midpoint_y = _p(keep, midpoint_y, np.float64)
    a.trace_index = _p(keep, trace_index, np.int32)
    a.n_traces = len(trace_index) if trace_index is not None else len(midpoint_x)
    a.nx, a.ny = len(x_axis), len(y_axis)
    a.max_aperture_m = float(max_aperture_m)
    out = np.zeros((a.nx, a.ny), np.int32)
    a.fold = out.ctypes.data
    lib.oracle_fold_tile(ctypes.byref(a), int(threads or os.cpu_count() or 1))
    return out


def curved_dsr(t0_s, h_s, h_r, v0, k) -> float:
    return _load().oracle_curved_dsr(t0_s, h_s, h_r, v0, k)


def vti_dsr(t0_s, h_s, h_r, v, eta) -> float:
    return _load().oracle_vti_dsr(t0_s, h_s, h_r, v, eta)


def vti_leg(tau, h, v, eta) -> float:
    return _load().oracle_vti_leg(tau, h, v, eta)


def integrate_traces(amp: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    amp = np.ascontiguousarray(amp, dtype=np.float32)
    ntr, ns = amp.shape
    p = np.zeros((ntr, ns + 1))
    s = np.zeros(ntr)
    _load().oracle_integrate_traces(amp.ctypes.data, ntr, ns, p.ctypes.data, s.ctypes.data)
    return p, s


__all__ = ["MigrateInputs", "OracleResult", "migrate", "build", "curved_dsr", "vti_dsr",
           "vti_leg", "integrate_traces", "AA_TRIANGLE", "ETA_3D"]


def triangle_at(trace, u: float, L: float) -> float:
    """Anti-aliased sample of ``trace`` (rounded to f32): triangle of half
    width L >= 1 at fractional position u, FP64."""
    tr = np.ascontiguousarray(trace, dtype=np.float32)
    return float(_load().oracle_triangle_at(tr.ctypes.data, int(tr.size), float(u), float(L)))


def interp_sample(trace, t: float, method: int) -> float:
    """One sample of ``trace`` (rounded to f32) at fractional position t with
    interpolation code ``method`` (PSTM_INTERP_*), FP64."""
    tr = np.ascontiguousarray(trace, dtype=np.float32)
    return float(_load().oracle_interp_sample(tr.ctypes.data, int(tr.size), float(t), int(method)))
