"""Migration of a trace set streamed from disk (Zarr v2 + Parquet headers).

The reference's executor reads each tile's traces from Zarr on the host,
with a prefetch thread for the next tile (pstm/pipeline/executor.py:
1203-1253, 1360-1395).  Here the trace array streams into HBM chunk by chunk
through pinned buffers (pm_amd.io.stream), and the job migrates group by
group (for offset-binned data, one common-offset bin per group): each group
is a plan over every output tile, launched as soon as its rows have arrived,
while the following groups are still being read and copied.  All groups
accumulate into one image and fold.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from pm_amd.io.stream import ChunkStreamer, StreamStats
from pm_amd.io.zarr_v2 import ZarrV2Array, read_headers


@dataclass
class StreamedResult:
    image: object          # (nx, ny, nt) f32, device
    fold: object           # (nx, ny) i32, device
    stats: StreamStats
    groups: int
    wall_s: float          # streaming + migration, end to end
    executor: object


def group_bounds(hdr: dict, key: str | None, n: int, rows_per_group: int) -> list[tuple[int, int]]:
    """Row ranges of the groups: runs of equal ``hdr[key]`` (the rows must be
    stored group by group), or fixed blocks of rows_per_group."""
    if key is not None and key in hdr:
        v = np.asarray(hdr[key])
        cut = np.flatnonzero(v[1:] != v[:-1]) + 1
        if len(np.unique(v)) != len(cut) + 1:
            raise ValueError(f"traces are not stored group by group in column {key!r}")
        edges = np.concatenate([[0], cut, [n]])
    else:
        edges = np.concatenate([np.arange(0, n, rows_per_group), [n]])
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]


def migrate_streamed(kernel, traces_path, headers_path, x_axis, y_axis, t_axis_ms, config, *, cube=None,
                     vrms_1d=None, eta_cube=None, tile: int = 64, group_key: str | None = "BIN",
                     rows_per_group: int = 100_000, traveltime: int | None = None, surrogate=None,
                     node_spacing: int = 16, n_buffers: int = 8, n_threads: int = 4,
                     region: tuple[int, int, int, int] | None = None) -> StreamedResult:
    """Stream a Zarr trace array to the GPU and migrate it, group by group.
    With a learned traveltime, every tile is checked first
    (TileExecutor.check_surrogate) and a tile that fails migrates with the
    exact one.  region = (x0, x1, y0, y1) pillars: migrate only the tiles
    that cover it (the whole image and fold are still allocated)."""
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import SyntheticJob

    arr = ZarrV2Array.open(traces_path)
    hdr = read_headers(headers_path)
    n, ns = arr.shape
    if len(hdr["SOU_X"]) != n:
        raise ValueError(f"{headers_path}: {len(hdr['SOU_X'])} headers for {n} traces")
    dt_ms = float(arr.attrs.get("sample_rate_ms", 0.0))
    if not dt_ms > 0:
        raise ValueError(f"{traces_path}: sample_rate_ms attribute missing")
    t0_ms = float(arr.attrs.get("start_time_ms", 0.0))
    dev = kernel.device
    dest = torch.empty((n, ns), dtype=torch.float32, device=dev)
    meta = {"eta_cube": eta_cube} if eta_cube is not None else {}
    job = SyntheticJob("streamed", dest, hdr["SOU_X"], hdr["SOU_Y"], hdr["REC_X"], hdr["REC_Y"], dt_ms, t0_ms,
                       np.asarray(x_axis, float), np.asarray(y_axis, float), np.asarray(t_axis_ms, float),
                       config, vrms_1d=vrms_1d, cube=cube, tile=tile, meta=meta)
    t_start = time.perf_counter()
    ex = TileExecutor(kernel, job, spatial_sort=False)   # rows stay in file order
    groups = group_bounds(hdr, group_key, n, rows_per_group)
    if region is not None:
        from pm_amd.autotune import region_tiles

        tiles = region_tiles(ex, *region)
    else:
        tiles = list(range(len(ex.tiles)))
    learned, exact = [], tiles
    if surrogate is not None:
        # geometry and the model only: the amplitudes have not arrived yet
        ex.check_surrogate(surrogate, node_spacing=node_spacing, tiles=tiles)
        learned = [i for i in tiles if ex.tiles[i].learned_ok]
        exact = [i for i in tiles if not ex.tiles[i].learned_ok]
    plans = []
    for g in groups:
        if learned:
            plans.append((g, ex.plan(row_range=g, tiles=learned, traveltime=traveltime, surrogate=surrogate,
                                     node_spacing=node_spacing)))
        if exact:
            plans.append((g, ex.plan(row_range=g, tiles=exact)))
    ex.reset()
    streamer = ChunkStreamer(arr, dest, n_buffers=n_buffers, n_threads=n_threads).start()
    try:
        for g, p in plans:
            streamer.wait_rows(*g)
            p.run(use_graph=False)
        torch.cuda.synchronize()
    finally:
        streamer.close()
        for _, p in plans:
            p.destroy()
    wall = time.perf_counter() - t_start
    img, fold = ex.assemble()
    return StreamedResult(img, fold, streamer.stats, len(groups), wall, ex)
