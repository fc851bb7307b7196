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
    eta: Any = None      # VTI: (tnx, tny, nt) f64, or None (constant eta)
    learned_ok: bool = False   # surrogate passed this tile's held-out check
    heldout_error: float | None = None   # max held-out traveltime error (samples)

    @property
    def n_pillars(self) -> int:
        return (self.x1 - self.x0) * (self.y1 - self.y0)


def learned_traveltime(surrogate) -> int:
    """The PSTM_TT_* code of a trained surrogate's traveltime."""
    from pm_amd.kernels import _abi

    return _abi.TT_CURVED_LEARNED if getattr(surrogate, "kind", "vti") == "curved" else _abi.TT_VTI_LEARNED


class TileExecutor:
    """Runs one migration job tile by tile on the GPU through HipKernel."""

    def __init__(self, kernel: HipKernel, job, config: KernelConfig | None = None,
                 spatial_sort: bool = True, time_windows=None):
        """``time_windows`` (pm_amd.time_variant.compute_time_windows over the
        job's output axis): migrate only the window samples; the images are
        then (tnx, tny, n_window_samples) and ``assemble_uniform`` resamples
        them onto the uniform axis."""
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
        amp = job.amplitudes if job.amplitudes.is_cuda else job.amplitudes.to(dev)
        if spatial_sort:
            amp = amp.index_select(0, torch.as_tensor(perm, device=dev)).contiguous()
        # else: the job's own tensor (a streaming destination is filled later)

        def d(a):
            return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)[perm]), dtype=f64,
                                   device=dev)

        self.traces = DeviceTraces(amp, d(job.source_x), d(job.source_y), d(job.receiver_x),
                                   d(job.receiver_y), d(mx), d(my), None, job.dt_ms,
                                   job.t_start_ms)
        self.time_windows = list(time_windows) if time_windows else None
        t_axis = np.asarray(job.t_axis_ms, dtype=np.float64)
        tsel = None
        if self.time_windows:
            from pm_amd.time_variant import sample_times, uniform_index

            t_axis = sample_times(self.time_windows)
            tsel = torch.as_tensor(uniform_index(t_axis, job.t_axis_ms), device=dev)
        self.t_axis = torch.as_tensor(t_axis, dtype=f64, device=dev)
        self.nt = len(t_axis)
        nx, ny = len(job.x_axis), len(job.y_axis)
        self.shape = (nx, ny, self.nt)

        def on_axis(a):
            # velocity / eta on the uniform axis, then at the window samples
            return a if tsel is None else a.index_select(-1, tsel).contiguous()

        if job.cube is not None:
            vfull = on_axis(job.cube.resample(job.x_axis, job.y_axis, job.t_axis_ms, device=dev))
        else:
            vfull = None
            v1 = on_axis(torch.as_tensor(job.vrms_1d, dtype=f64, device=dev))
        ecube = job.meta.get("eta_cube") if self.config.vti_enabled else None
        efull = (on_axis(ecube.resample(job.x_axis, job.y_axis, job.t_axis_ms, device=dev))
                 if ecube is not None else None)
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
                torch.zeros((tnx, tny), dtype=torch.int32, device=dev),
                efull[x0:x1, y0:y1].contiguous() if efull is not None else None))
        del vfull, efull

    def reset(self) -> None:
        self.image_storage.zero_()
        for t in self.tiles:
            t.fold.zero_()

    def run(self, count_effective: bool = False, surrogate=None, node_spacing: int = 16,
            tiles: list[int] | None = None) -> int | None:
        """One pass over every tile (or the listed ones), asynchronous unless
        counting.  With a surrogate (VTI jobs), tiles whose held-out check
        passed (``check_surrogate``) use the learned traveltime and the others
        the exact one."""
        from pm_amd.kernels import _abi

        if surrogate is not None and self.time_windows:
            raise ValueError("the learned traveltime needs a uniform output axis (no time windows)")
        total = 0
        for i in (range(len(self.tiles)) if tiles is None else tiles):
            t = self.tiles[i]
            tt = None
            if surrogate is not None and t.learned_ok:
                tt = learned_traveltime(surrogate)
            e = self.kernel.migrate_device(self.traces, t.image, t.fold, t.x_axis, t.y_axis,
                                           self.t_axis, t.vrms, self.config, trace_index=t.rows,
                                           count_effective=count_effective, eta=t.eta, traveltime=tt,
                                           surrogate=surrogate, node_spacing=node_spacing)
            if count_effective:
                total += int(e)
        return total if count_effective else None

    def plan(self, tiles: list[int] | None = None, traveltime: int | None = None, surrogate=None,
             node_spacing: int = 16, count_effective: bool = False, row_range: tuple[int, int] | None = None):
        """A device-resident work queue over the listed tiles (all by default):
        HipPlan.run() migrates them in one persistent, hipGraph-captured gather
        launch, heaviest tiles first.  Trace lists may hold the rows of any
        number of common-offset bins; all of them land in the shared image.
        ``row_range`` (r0, r1) keeps only the executor rows in [r0, r1) (one
        group of a streamed job)."""
        if surrogate is not None and self.time_windows:
            raise ValueError("the learned traveltime needs a uniform output axis (no time windows)")
        sel = range(len(self.tiles)) if tiles is None else tiles
        descs = []
        for t in (self.tiles[i] for i in sel):
            rows = t.rows
            if row_range is not None:
                rows = rows[(rows >= row_range[0]) & (rows < row_range[1])].contiguous()
            descs.append(dict(image=t.image, fold=t.fold, x_axis=t.x_axis, y_axis=t.y_axis, vrms=t.vrms,
                              trace_index=rows, eta=t.eta))
        return self.kernel.create_plan(self.traces, descs, self.t_axis, self.config, traveltime=traveltime,
                                       count_effective=count_effective, surrogate=surrogate,
                                       node_spacing=node_spacing)

    def check_surrogate(self, surrogate, node_spacing: int = 16, n: int = 8192, bar: float = 0.25,
                        seed: int = 0, tiles: list[int] | None = None) -> list[float]:
        """Per-tile held-out check of the learned traveltime: n random
        (pillar, in-aperture trace, output sample) triples of the tile, the
        learned time as the gather forms it (pm_amd.surrogate.lattice_times)
        against the exact VTI time, in input samples.  A tile passes when the
        maximum error is at most ``bar``; the others fall back to the exact
        traveltime.  Returns the per-tile maxima."""
        import torch

        from pm_amd.surrogate import (curved_leg_times, curved_velocity_at, lattice_times, pillar_group_ends,
                                      vti_leg_times)

        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        tr = self.traces
        tau_axis = self.t_axis / 2000.0
        dt_in = tr.sample_rate_ms / 1000.0
        ns = tr.n_samples
        out = []
        for i in (range(len(self.tiles)) if tiles is None else tiles):
            t = self.tiles[i]
            tnx, tny = t.x1 - t.x0, t.y1 - t.y0
            m = 4 * n
            pil = torch.randint(0, tnx * tny, (m,), device=self.device, generator=g)
            row = t.rows[torch.randint(0, int(t.rows.shape[0]), (m,), device=self.device, generator=g)].long()
            k = torch.randint(0, self.nt, (m,), device=self.device, generator=g)
            ox, oy = t.x_axis[pil // tny], t.y_axis[pil % tny]
            keep = ((ox - tr.midpoint_x[row]) ** 2 + (oy - tr.midpoint_y[row]) ** 2).sqrt() <= self.config.max_aperture_m
            pil, row, k, ox, oy = pil[keep][:n], row[keep][:n], k[keep][:n], ox[keep][:n], oy[keep][:n]
            ds = ((ox - tr.source_x[row]) ** 2 + (oy - tr.source_y[row]) ** 2).sqrt()
            dr = ((ox - tr.receiver_x[row]) ** 2 + (oy - tr.receiver_y[row]) ** 2).sqrt()
            # the gather evaluates the network at three pillars of each
            # pillar's group and interpolates between them
            pe, w = pillar_group_ends(pil, tnx * tny)
            ex_, ey_ = t.x_axis[pe // tny], t.y_axis[pe % tny]
            ds_e = ((ex_ - tr.source_x[row][:, None]) ** 2 + (ey_ - tr.source_y[row][:, None]) ** 2).sqrt()
            dr_e = ((ex_ - tr.receiver_x[row][:, None]) ** 2 + (ey_ - tr.receiver_y[row][:, None]) ** 2).sqrt()
            if getattr(surrogate, "kind", "vti") == "curved":
                if float(surrogate.curved_v0) != float(self.config.curved_ray_v0):
                    raise ValueError("the curved surrogate was trained for another surface velocity")
                # v(z) at the image depth of each sample, the gradient k as the fifth input
                v0, kg = float(self.config.curved_ray_v0), float(self.config.curved_ray_k)
                vg_axis = (v0 * curved_velocity_at(tau_axis, v0, kg)).sqrt()
                v_at = lambda kk: vg_axis[kk]  # noqa: E731
                e_at = lambda kk: torch.full(kk.shape, kg, dtype=torch.float64, device=self.device)  # noqa: E731
                ends = {"ds": ds_e, "dr": dr_e, "v_at": lambda kk, e: v_at(kk), "eta_at": lambda kk, e: e_at(kk),
                        "w": w}
                learned = lattice_times(surrogate, tau_axis, k, ds, dr, v_at, e_at, node_spacing, ends)
                tau = tau_axis[k]
                exact = curved_leg_times(tau, ds, v0, kg) + curved_leg_times(tau, dr, v0, kg)
            else:
                vflat = t.vrms.reshape(-1, self.nt) if t.vrms.dim() == 3 else None
                eflat = t.eta.reshape(-1, self.nt) if t.eta is not None else None
                v_at = (lambda kk: vflat[pil, kk]) if vflat is not None else (lambda kk: t.vrms[kk])
                if eflat is not None:
                    e_at = lambda kk: eflat[pil, kk]  # noqa: E731
                else:
                    e_at = lambda kk: torch.full(kk.shape, float(self.config.vti_eta_constant),  # noqa: E731
                                                 dtype=torch.float64, device=self.device)
                v_e = (lambda kk, e: vflat[pe[:, e], kk]) if vflat is not None else (lambda kk, e: t.vrms[kk])
                e_e = (lambda kk, e: eflat[pe[:, e], kk]) if eflat is not None else (lambda kk, e: e_at(kk))
                ends = {"ds": ds_e, "dr": dr_e, "v_at": v_e, "eta_at": e_e, "w": w}
                learned = lattice_times(surrogate, tau_axis, k, ds, dr, v_at, e_at, node_spacing, ends)
                tau = tau_axis[k]
                exact = vti_leg_times(tau, ds, v_at(k), e_at(k)) + vti_leg_times(tau, dr, v_at(k), e_at(k))
            inside = exact / dt_in < ns - 1
            err = float(((learned - exact).abs() / dt_in)[inside].max()) if bool(inside.any()) else 0.0
            t.heldout_error = err
            t.learned_ok = err <= bar
            out.append(err)
        return out

    def output_grid(self):
        """The job's output axes and bin sizes (KernelConfig.grid_dx / dy)."""
        from pm_amd.pipeline.output import OutputGrid

        return OutputGrid.from_axes(self.job.x_axis, self.job.y_axis, self.job.t_axis_ms, self.config.grid_dx,
                                    self.config.grid_dy)

    def header_sums(self, sums=None, rows_of=None):
        """Per-output-bin trace-header sums of the whole job: every tile adds
        its trace query, binned by midpoint and clipped to the tile, as the
        reference's executor does after each tile (executor.py:1783-1806).
        ``rows_of(i)`` overrides tile i's rows (a gather bin's selection).
        Asynchronous; returns the (new, or given) HeaderSums."""
        from pm_amd.pipeline.output import HeaderSums

        nx, ny, _ = self.shape
        sums = sums if sums is not None else HeaderSums.zeros(nx, ny, self.device)
        dx, dy = float(self.config.grid_dx), float(self.config.grid_dy)
        for i, t in enumerate(self.tiles):
            rows = t.rows if rows_of is None else rows_of(i)
            self.kernel.bin_headers(self.traces, rows, float(self.job.x_axis[t.x0]), float(self.job.y_axis[t.y0]),
                                    dx, dy, t.x0, t.y0, t.x1 - t.x0, t.y1 - t.y0, sums)
        return sums

    def finalize(self, out_dir, headers=None, **kw):
        """Write the stacked image (on the uniform axis, also with time
        windows), fold volume and bin headers
        (pm_amd.pipeline.output.finalize_stack); computes the header sums when
        none are given.  Returns the WriteStats."""
        from pm_amd.pipeline.output import finalize_stack

        headers = headers if headers is not None else self.header_sums()
        image, fold = self.assemble()
        # with time windows: the window samples back on the uniform axis, as
        # the reference resamples each time-variant tile (executor.py:1720-1726)
        image = self.to_uniform(image)
        return finalize_stack(out_dir, image, fold, self.output_grid(), headers, **kw)

    def gathers(self, bins) -> "GatherSet":
        """Gather mode: one image, fold and header sums per GatherBin, each
        tile's trace list narrowed to the bin's traces (executor.py:960-975,
        1533-1590)."""
        return GatherSet(self, list(bins))

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

    def assemble_uniform(self):
        """(nx, ny, nt_uniform) f32 image on the job's uniform axis: the
        window samples resampled as the reference does after a time-variant
        tile (pstm/pipeline/executor.py:1720-1726; np.interp onto
        k * dt, k < nt), or ``assemble()``'s image without time windows."""
        img, _ = self.assemble()
        return self.to_uniform(img)

    def to_uniform(self, image):
        """An (nx, ny, n) image of this executor's compute axis on the job's
        uniform axis (unchanged without time windows)."""
        from pm_amd.time_variant import sample_times, uniform_times

        if not self.time_windows:
            return image
        t = np.asarray(self.job.t_axis_ms, dtype=np.float64)
        return self.kernel.resample_uniform(image.contiguous(), sample_times(self.time_windows),
                                            uniform_times(t[1] - t[0], len(t)))

    def host_traces(self, rows: np.ndarray | None = None) -> dict:
        """Host copies (in the executor's row order) for the CPU oracle."""
        tr = self.traces
        sel = slice(None) if rows is None else rows
        out = {"amplitudes": tr.amplitudes[sel].cpu().numpy()}
        for name in ("source_x", "source_y", "receiver_x", "receiver_y", "midpoint_x", "midpoint_y"):
            out[name] = getattr(tr, name)[sel].cpu().numpy()
        return out


class GatherSet:
    """Per-bin output volumes of a gather-mode job.  The reference migrates
    each bin in its own pass over the tiles into its own memmaps
    (pstm/pipeline/executor.py:960-975, 838-860); here every (tile, bin) pair
    is one entry of a single plan, so one persistent launch fills all bins'
    images, and the header sums of each bin are accumulated on the device."""

    def __init__(self, ex: TileExecutor, bins: list):
        import torch

        if not bins:
            raise ValueError("gather mode needs at least one bin")
        self.ex = ex
        self.bins = bins
        dev = ex.device
        slab = int(ex.image_storage.shape[0])
        self.images = torch.zeros((len(bins), slab), dtype=torch.float32, device=dev)
        self.folds = [[torch.zeros((t.x1 - t.x0, t.y1 - t.y0), dtype=torch.int32, device=dev) for t in ex.tiles]
                      for _ in bins]
        self.rows = [[ex.kernel.select_rows(ex.traces, t.rows, gb) for t in ex.tiles] for gb in bins]
        self._views = []
        for b in range(len(bins)):
            off, views = 0, []
            for t in ex.tiles:
                n = (t.x1 - t.x0) * (t.y1 - t.y0) * ex.nt
                views.append(self.images[b, off:off + n].view(t.x1 - t.x0, t.y1 - t.y0, ex.nt))
                off += n
            self._views.append(views)
        self.headers = [ex.header_sums(rows_of=lambda i, b=b: self.rows[b][i]) for b in range(len(bins))]

    def reset(self) -> None:
        self.images.zero_()
        for fb in self.folds:
            for f in fb:
                f.zero_()

    def plan(self, **kw):
        """One HipPlan over every (tile, bin) pair with traces."""
        ex = self.ex
        descs = []
        for b in range(len(self.bins)):
            for i, t in enumerate(ex.tiles):
                rows = self.rows[b][i]
                if int(rows.shape[0]) == 0:
                    continue
                descs.append(dict(image=self._views[b][i], fold=self.folds[b][i], x_axis=t.x_axis, y_axis=t.y_axis,
                                  vrms=t.vrms, trace_index=rows, eta=t.eta))
        return ex.kernel.create_plan(ex.traces, descs, ex.t_axis, ex.config, **kw)

    def assemble(self, b: int):
        """Bin b's (nx, ny, nt) f32 image and (nx, ny) i32 fold on the device."""
        import torch

        nx, ny, nt = self.ex.shape
        img = torch.empty((nx, ny, nt), dtype=torch.float32, device=self.ex.device)
        fold = torch.empty((nx, ny), dtype=torch.int32, device=self.ex.device)
        for t, v, f in zip(self.ex.tiles, self._views[b], self.folds[b]):
            img[t.x0:t.x1, t.y0:t.y1] = v
            fold[t.x0:t.x1, t.y0:t.y1] = f
        return img, fold

    def finalize(self, out_dir, **kw):
        """Write every bin (pm_amd.pipeline.output.finalize_gathers)."""
        from pm_amd.pipeline.output import finalize_gathers

        def volumes(b):
            img, fold = self.assemble(b)
            return self.ex.to_uniform(img), fold, self.headers[b]

        return finalize_gathers(out_dir, self.ex.kernel, self.bins, volumes, self.ex.output_grid(), **kw)
