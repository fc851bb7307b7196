"""The executor's output step: per-bin trace-header sums, gather bins, and the
finalize write-out.

Reference behaviour restated here:
  * after each tile, per-output-bin sums of trace count, offset and azimuth
    sine / cosine over the tile's trace query, midpoints binned with
    floor((m - tile_min) / d) clipped to the tile
    (pstm/pipeline/executor.py:1783-1806) — ``HeaderSums``, filled on the
    device by ``HipKernel.bin_headers``;
  * gather mode: one image, fold and set of header sums per offset or
    offset-vector-tile bin, each fed only the traces that pass the bin's
    closed-interval filter (executor.py:1537-1554; GatherBin,
    pstm/config/models.py:677-760) — ``GatherBin``, ``HipKernel.select_rows``;
  * finalize of the stack (executor.py:2032-2123): ``migrated_stack.zarr``
    (f32, chunks (chunk_x, chunk_y, chunk_t or nt), grid attributes),
    ``fold.zarr`` (i32, the 3-D fold: the pillar fold repeated along t) and
    ``bin_headers.parquet`` (executor.py:2295-2329) — ``finalize_stack``;
  * finalize of gathers (executor.py:2125-2293): per bin
    ``gathers/gather_bin_NNN.zarr`` (image / fold where fold > 0, else 0),
    ``gathers/fold_bin_NNN.zarr``, then ``gather_headers.parquet`` and
    ``gathers_index.parquet`` — ``finalize_gathers``.

Differences, on purpose: chunks are stored uncompressed (``compressor:
null``; any Zarr v2 reader reads them, and this package has no Blosc codec);
the fold volumes are chunked like the images (the reference leaves their
chunking to zarr); the image is accumulated in f32 where the reference's
memmap is f64 that it casts to f32 on write.

Volumes go from HBM to disk one x-slab of ``chunk_x`` pillars at a time:
the slab is copied into pinned host memory on a side stream while a thread
pool writes the previous slab's chunk files.
"""

from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np

from pm_amd.io.zarr_v2 import ZarrV2Writer


@dataclass(frozen=True)
class GatherBin:
    """One output gather volume: an absolute-offset range or an offset-vector
    tile (signed o_x = r_x - s_x, o_y = r_y - s_y ranges); bounds are closed and
    a bound of None is not checked (pstm/config/models.py:677-760)."""

    bin_type: str                      # "offset" or "ovt"
    name: str | None = None
    offset_min: float | None = None
    offset_max: float | None = None
    ovt_x_min: float | None = None
    ovt_x_max: float | None = None
    ovt_y_min: float | None = None
    ovt_y_max: float | None = None

    def __post_init__(self):
        if self.bin_type not in ("offset", "ovt"):
            raise ValueError(f"bin_type must be 'offset' or 'ovt', not {self.bin_type!r}")

    @classmethod
    def offset_bin(cls, offset_min: float, offset_max: float, name: str | None = None) -> "GatherBin":
        return cls("offset", name, offset_min=offset_min, offset_max=offset_max)

    @classmethod
    def ovt_bin(cls, x_min: float, x_max: float, y_min: float, y_max: float, name: str | None = None) -> "GatherBin":
        return cls("ovt", name, ovt_x_min=x_min, ovt_x_max=x_max, ovt_y_min=y_min, ovt_y_max=y_max)

    @property
    def label(self) -> str:
        """The bin's name, or the reference's default name (executor.py:2140-2146)."""
        if self.name:
            return self.name
        if self.bin_type == "offset":
            return f"offset_{self.offset_min:.0f}_{self.offset_max:.0f}"
        return f"ovt_x{self.ovt_x_min:.0f}_{self.ovt_x_max:.0f}_y{self.ovt_y_min:.0f}_{self.ovt_y_max:.0f}"

    def type_fields(self) -> dict:
        """Bin-type specific attributes (executor.py:2188-2197)."""
        if self.bin_type == "offset":
            return {"offset_min": self.offset_min, "offset_max": self.offset_max,
                    "offset_center": (self.offset_min + self.offset_max) / 2}
        return {"ovt_x_min": self.ovt_x_min, "ovt_x_max": self.ovt_x_max, "ovt_y_min": self.ovt_y_min,
                "ovt_y_max": self.ovt_y_max}


@dataclass
class HeaderSums:
    """Per-output-bin trace-header sums over the (nx, ny) grid: trace count
    (i32) and f64 sums of offset, sin(azimuth) and cos(azimuth)."""

    trace_count: Any
    offset_sum: Any
    azimuth_sin_sum: Any
    azimuth_cos_sum: Any

    @classmethod
    def zeros(cls, nx: int, ny: int, device) -> "HeaderSums":
        import torch

        f64 = dict(dtype=torch.float64, device=device)
        return cls(torch.zeros((nx, ny), dtype=torch.int32, device=device), torch.zeros((nx, ny), **f64),
                   torch.zeros((nx, ny), **f64), torch.zeros((nx, ny), **f64))

    @property
    def shape(self) -> tuple[int, int]:
        return tuple(int(s) for s in self.trace_count.shape)

    def to_host(self) -> dict:
        def h(a):
            return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)

        return {"trace_count": h(self.trace_count), "offset_sum": h(self.offset_sum),
                "azimuth_sin_sum": h(self.azimuth_sin_sum), "azimuth_cos_sum": h(self.azimuth_cos_sum)}

    def averages(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(trace_count, offset_avg, azimuth_avg in degrees [0, 360)), 0 where
        a bin has no trace (executor.py:2107-2113)."""
        h = self.to_host()
        n = h["trace_count"]
        with np.errstate(invalid="ignore", divide="ignore"):
            off = np.where(n > 0, h["offset_sum"] / n, 0.0)
            az = np.degrees(np.arctan2(h["azimuth_sin_sum"], h["azimuth_cos_sum"])) % 360
            az = np.where(n > 0, az, 0.0)
        return n, off, az


@dataclass(frozen=True)
class OutputGrid:
    """Output axes and bin sizes, for the volumes' attributes and the header
    tables (OutputGrid.get_output_coordinates, pstm/config/output_grid.py:
    582-632: an axis-aligned grid)."""

    x: np.ndarray
    y: np.ndarray
    t_ms: np.ndarray
    dx: float
    dy: float
    dt_ms: float

    @classmethod
    def from_axes(cls, x, y, t_ms, dx: float | None = None, dy: float | None = None) -> "OutputGrid":
        x, y, t = (np.asarray(a, dtype=np.float64) for a in (x, y, t_ms))
        step = lambda a: float(a[1] - a[0]) if len(a) > 1 else 0.0  # noqa: E731
        return cls(x, y, t, float(dx) if dx is not None else step(x), float(dy) if dy is not None else step(y),
                   step(t))

    @property
    def shape(self) -> tuple[int, int, int]:
        return len(self.x), len(self.y), len(self.t_ms)

    def attrs(self) -> dict:
        """The stack volume's attributes (executor.py:2082-2092)."""
        return {"x_min": float(self.x[0]), "x_max": float(self.x[-1]), "dx": self.dx,
                "y_min": float(self.y[0]), "y_max": float(self.y[-1]), "dy": self.dy,
                "t_min_ms": float(self.t_ms[0]), "t_max_ms": float(self.t_ms[-1]), "dt_ms": self.dt_ms,
                "migration_type": "PSTM"}

    def header_columns(self) -> dict:
        """ix, iy, x, y per output bin, (nx, ny) order (executor.py:2310-2318)."""
        ix, iy = np.meshgrid(np.arange(len(self.x)), np.arange(len(self.y)), indexing="ij")
        X, Y = np.meshgrid(self.x, self.y, indexing="ij")
        return {"ix": ix.ravel().astype(np.int32), "iy": iy.ravel().astype(np.int32),
                "x": X.ravel().astype(np.float64), "y": Y.ravel().astype(np.float64)}


@dataclass
class WriteStats:
    """Bytes written and wall time of a finalize step."""

    bytes: int = 0
    seconds: float = 0.0
    files: list = field(default_factory=list)

    @property
    def gb_per_s(self) -> float:
        return self.bytes / self.seconds / 1e9 if self.seconds > 0 else 0.0

    def add(self, other: "WriteStats") -> None:
        self.bytes += other.bytes
        self.seconds += other.seconds
        self.files.extend(other.files)


def _chunks(shape, chunk_x: int, chunk_y: int, chunk_t: int | None) -> tuple[int, int, int]:
    nx, ny, nt = shape
    return max(1, min(chunk_x, nx)), max(1, min(chunk_y, ny)), max(1, min(chunk_t or nt, nt))


def write_device_volume(path, volume, chunks, attrs: dict | None = None, workers: int = 4) -> WriteStats:
    """Write a (nx, ny, nt) float32 device tensor as a Zarr v2 array: one
    x-slab of chunk_x pillars at a time into pinned host memory on a side
    stream (after the work queued on the current stream), chunk files written
    by ``workers`` threads while the next slab copies."""
    import torch

    t0 = time.perf_counter()
    if volume.dim() != 3 or volume.dtype != torch.float32 or not volume.is_contiguous():
        raise ValueError("volume must be a contiguous (nx, ny, nt) float32 tensor")
    nx, ny, nt = (int(s) for s in volume.shape)
    w = ZarrV2Writer(path, (nx, ny, nt), chunks, np.float32, attrs)
    cx = w.chunks[0]
    dev = volume.device
    bufs = [torch.empty((cx, ny, nt), dtype=torch.float32, pin_memory=True) for _ in range(2)]
    pending: list[list] = [[], []]
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    written = 0
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for i in range(w.grid[0]):
            b = i % 2
            for f in pending[b]:
                written += f.result()
            x0, x1 = i * cx, min((i + 1) * cx, nx)
            with torch.cuda.stream(side):
                bufs[b][: x1 - x0].copy_(volume[x0:x1], non_blocking=True)
            side.synchronize()
            host = bufs[b].numpy()[: x1 - x0]
            pending[b] = [pool.submit(w.write_chunk, (i, j, k), host[:, w.bounds((i, j, k))[1], w.bounds((i, j, k))[2]])
                          for j in range(w.grid[1]) for k in range(w.grid[2])]
        for p in pending:
            for f in p:
                written += f.result()
    return WriteStats(written, time.perf_counter() - t0, [str(path)])


def write_fold_volume(path, fold2d, nt: int, chunks, attrs: dict | None = None, workers: int = 4) -> WriteStats:
    """The 3-D fold (the (nx, ny) pillar fold repeated along t) as an int32
    Zarr v2 array; the repetition happens per chunk on the host."""
    t0 = time.perf_counter()
    f = fold2d.detach().cpu().numpy() if hasattr(fold2d, "detach") else np.asarray(fold2d)
    f = np.ascontiguousarray(f, dtype=np.int32)
    nx, ny = f.shape
    w = ZarrV2Writer(path, (nx, ny, nt), chunks, np.int32, attrs)

    def one(idx):
        bx, by, bt = w.bounds(idx)
        return w.write_chunk(idx, np.broadcast_to(f[bx, by, None], (bx.stop - bx.start, by.stop - by.start,
                                                                      bt.stop - bt.start)))

    with ThreadPoolExecutor(max_workers=workers) as pool:
        written = sum(pool.map(one, np.ndindex(*w.grid)))
    return WriteStats(written, time.perf_counter() - t0, [str(path)])


def _write_parquet(path, cols: dict) -> None:
    import pyarrow as pa
    import pyarrow.parquet as pq

    pq.write_table(pa.table(cols), str(path))


def header_table(grid: OutputGrid, sums: HeaderSums) -> dict:
    """Columns of the per-bin header table, rows with trace_count > 0
    (executor.py:2295-2324)."""
    n, off, az = sums.averages()
    cols = grid.header_columns()
    cols.update({"trace_count": n.ravel().astype(np.int32), "offset_avg": off.ravel().astype(np.float32),
                 "azimuth_avg": az.ravel().astype(np.float32)})
    keep = cols["trace_count"] > 0
    return {k: v[keep] for k, v in cols.items()}


def finalize_stack(out_dir, image, fold2d, grid: OutputGrid, headers: HeaderSums | None = None, *,
                   chunk_x: int = 64, chunk_y: int = 64, chunk_t: int | None = None,
                   fold_volume: bool = True) -> WriteStats:
    """Write the stacked image (device (nx, ny, nt) f32; the Kirchhoff sum, no
    fold normalisation, executor.py:2054-2060), the fold volume and the bin
    headers into ``out_dir`` as the reference's stack finalize does."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    if tuple(image.shape) != grid.shape:
        raise ValueError(f"image {tuple(image.shape)} != grid {grid.shape}")
    ch = _chunks(grid.shape, chunk_x, chunk_y, chunk_t)
    st = write_device_volume(out / "migrated_stack.zarr", image, ch, grid.attrs())
    if fold_volume:
        st.add(write_fold_volume(out / "fold.zarr", fold2d, grid.shape[2], ch,
                                 {"description": "Migration fold map"}))
    if headers is not None:
        t0 = time.perf_counter()
        _write_parquet(out / "bin_headers.parquet", header_table(grid, headers))
        st.add(WriteStats(0, time.perf_counter() - t0, [str(out / "bin_headers.parquet")]))
    return st


def finalize_gathers(out_dir, kernel, bins: list[GatherBin], volumes, grid: OutputGrid, *, chunk_x: int = 64,
                     chunk_y: int = 64, chunk_t: int | None = None) -> WriteStats:
    """Write every gather bin as the reference's gather finalize does.
    ``volumes(b)`` returns bin b's (image (nx, ny, nt) f32, fold (nx, ny) i32,
    HeaderSums) on the device; the image is divided by the fold on the device
    (``kernel.fold_normalize``) before it is written."""
    import pyarrow as pa

    out = Path(out_dir)
    gdir = out / "gathers"
    gdir.mkdir(parents=True, exist_ok=True)
    ch = _chunks(grid.shape, chunk_x, chunk_y, chunk_t)
    st = WriteStats()
    tables, index = [], []
    for bid, gb in enumerate(bins):
        image, fold2d, sums = volumes(bid)
        norm = kernel.fold_normalize(image, fold2d)
        attrs = grid.attrs()
        attrs.update({"gather_bin_id": bid, "gather_bin_type": gb.bin_type, "gather_bin_name": gb.label})
        attrs.update(gb.type_fields())
        st.add(write_device_volume(gdir / f"gather_bin_{bid:03d}.zarr", norm, ch, attrs))
        del norm
        st.add(write_fold_volume(gdir / f"fold_bin_{bid:03d}.zarr", fold2d, grid.shape[2], ch,
                                 {"gather_bin_id": bid}))
        cols = header_table(grid, sums)
        m = len(cols["ix"])
        cols.update({"gather_bin_id": np.full(m, bid, np.int32), "gather_bin_type": [gb.bin_type] * m,
                     "gather_bin_name": [gb.label] * m})
        for k, v in gb.type_fields().items():
            # offset_min -> offset_bin_min etc. (executor.py:2239-2247)
            col = "offset_bin_" + k[len("offset_"):] if gb.bin_type == "offset" else k
            cols[col] = np.full(m, v, np.float32)
        tables.append(pa.table(cols))
        rec = {"bin_id": bid, "bin_type": gb.bin_type, "bin_name": gb.label,
               "volume_path": f"gathers/gather_bin_{bid:03d}.zarr"}
        rec.update(gb.type_fields())
        index.append(rec)
    t0 = time.perf_counter()
    import pyarrow.parquet as pq

    if tables:
        pq.write_table(pa.concat_tables(tables, promote_options="default"), str(out / "gather_headers.parquet"))
    keys = list(dict.fromkeys(k for rec in index for k in rec))   # union, first-seen order
    pq.write_table(pa.table({k: [rec.get(k) for rec in index] for k in keys}), str(out / "gathers_index.parquet"))
    st.add(WriteStats(0, time.perf_counter() - t0,
                      [str(out / "gather_headers.parquet"), str(out / "gathers_index.parquet")]))
    return st
