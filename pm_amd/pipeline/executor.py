"""Device-resident tile executor.

The reference's executor loops over output tiles on the host: per tile it
queries the traces in the aperture-padded rectangle, builds a TraceBlock and
a zeroed OutputTile, fetches the tile's velocity slice, calls
kernel.migrate_tile and adds the tile into the global image and fold memmaps
(pstm/pipeline/executor.py:1397-1858, dispatch at 1691-1745, accumulation at
1776-1777).  Here the whole job lives in HBM (288 GB holds the traces, the
f64 velocity volume and the f32 image many times over): traces are uploaded
once in Z-order, every tile gets its row list, velocity slice and image slab
once, and a pass over the tiles is a sequence of asynchronous launches on one
stream with no host round trip.  The image is stored tile-major, (tile, ix,
iy, t), so that each launch writes one contiguous slab; ``assemble`` returns
the (nx, ny, nt) volume.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from pm_amd.kernels.base import KernelConfig
from pm_amd.kernels.hip_backend import DeviceTraces, HipKernel
from pm_amd.pipeline.spatial import morton_order, query_traces_for_tile


@dataclass
class DeviceTile:
    x0: int
    x1: int
    y0: int
    y1: int
    x_axis: Any          # (tnx,) f64
    y_axis: Any          # (tny,) f64
    vrms: Any            # (tnx, tny, nt) f64 or (nt,) f64
    rows: Any            # (n,) int32 trace rows in the aperture-padded rectangle
    image: Any           # (tnx, tny, nt) f32 (slab of the tile-major image)
    fold: Any            # (tnx, tny) i32

    @property
    def n_pillars(self) -> int:
        return (self.x1 - self.x0) * (self.y1 - self.y0)


class TileExecutor:
    """Runs one migration job tile by tile on the GPU through HipKernel."""

    def __init__(self, kernel: HipKernel, job, config: KernelConfig | None = None,
                 spatial_sort: bool = True):
        import torch

        self.kernel = kernel
        self.job = job
        self.config = config or job.config
        dev = kernel.device
        self.device = dev
        f64 = torch.float64
        mx, my = job.midpoint_x, job.midpoint_y
        perm = morton_order(mx, my) if spatial_sort else np.arange(len(mx))
        self.perm = perm
        pt = torch.as_tensor(perm, device=dev)
        amp = job.amplitudes if job.amplitudes.is_cuda else job.amplitudes.to(dev)
        amp = amp.index_select(0, pt).contiguous()

        def d(a):
            return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)[perm]), dtype=f64,
                                   device=dev)

        self.traces = DeviceTraces(amp, d(job.source_x), d(job.source_y), d(job.receiver_x),
                                   d(job.receiver_y), d(mx), d(my), None, job.dt_ms,
                                   job.t_start_ms)
        self.t_axis = torch.as_tensor(job.t_axis_ms, dtype=f64, device=dev)
        self.nt = len(job.t_axis_ms)
        nx, ny = len(job.x_axis), len(job.y_axis)
        self.shape = (nx, ny, self.nt)
        if job.cube is not None:
            vfull = job.cube.resample(job.x_axis, job.y_axis, job.t_axis_ms, device=dev)
        else:
            vfull = None
            v1 = torch.as_tensor(job.vrms_1d, dtype=f64, device=dev)
        tiles = job.tiles()
        slab = sum((x1 - x0) * (y1 - y0) for x0, x1, y0, y1 in tiles) * self.nt
        self.image_storage = torch.zeros(slab, dtype=torch.float32, device=dev)
        self.tiles: list[DeviceTile] = []
        off = 0
        ap = float(self.config.max_aperture_m)
        mx_d, my_d = self.traces.midpoint_x, self.traces.midpoint_y
        for x0, x1, y0, y1 in tiles:
            tnx, tny = x1 - x0, y1 - y0
            img = self.image_storage[off:off + tnx * tny * self.nt].view(tnx, tny, self.nt)
            off += tnx * tny * self.nt
            rows = query_traces_for_tile(mx_d, my_d, float(job.x_axis[x0]), float(job.x_axis[x1 - 1]),
                                         float(job.y_axis[y0]), float(job.y_axis[y1 - 1]), ap)
            v = vfull[x0:x1, y0:y1].contiguous() if vfull is not None else v1
            self.tiles.append(DeviceTile(
                x0, x1, y0, y1, torch.as_tensor(job.x_axis[x0:x1], dtype=f64, device=dev),
                torch.as_tensor(job.y_axis[y0:y1], dtype=f64, device=dev), v, rows, img,
                torch.zeros((tnx, tny), dtype=torch.int32, device=dev)))
        del vfull

    def reset(self) -> None:
        self.image_storage.zero_()
        for t in self.tiles:
            t.fold.zero_()

    def run(self, count_effective: bool = False) -> int | None:
        """One pass over every tile (asynchronous unless counting)."""
        total = 0
        for t in self.tiles:
            e = self.kernel.migrate_device(self.traces, t.image, t.fold, t.x_axis, t.y_axis,
                                           self.t_axis, t.vrms, self.config, trace_index=t.rows,
                                           count_effective=count_effective)
            if count_effective:
                total += int(e)
        return total if count_effective else None

    def contributions(self) -> int:
        """C = sum of the 3-D fold = sum over pillars of fold2d * nt."""
        return int(sum(int(t.fold.sum()) for t in self.tiles)) * self.nt

    def assemble(self):
        """(nx, ny, nt) f32 image and (nx, ny) i32 fold on the device."""
        import torch

        nx, ny, nt = self.shape
        img = torch.empty((nx, ny, nt), dtype=torch.float32, device=self.device)
        fold = torch.empty((nx, ny), dtype=torch.int32, device=self.device)
        for t in self.tiles:
            img[t.x0:t.x1, t.y0:t.y1] = t.image
            fold[t.x0:t.x1, t.y0:t.y1] = t.fold
        return img, fold

    def host_traces(self, rows: np.ndarray | None = None) -> dict:
        """Host copies (in the executor's row order) for the CPU oracle."""
        tr = self.traces
        sel = slice(None) if rows is None else rows
        out = {"amplitudes": tr.amplitudes[sel].cpu().numpy()}
        for name in ("source_x", "source_y", "receiver_x", "receiver_y", "midpoint_x", "midpoint_y"):
            out[name] = getattr(tr, name)[sel].cpu().numpy()
        return out
