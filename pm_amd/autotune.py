"""Launch-shape autotuner for the gather kernel.

The reference sizes tiles with a memory heuristic (pstm/pipeline/
tile_planner.py:26-88) and never measures.  Here the shape of one launch —
output samples per workgroup (256 x MW, MW sample blocks of 64 per wave) and
traces per LDS batch, pipelined staging on or off — is chosen by timing
candidate shapes as one plan over the job's busiest tiles with HIP events on
the launch stream, and the winner is set for every later launch of the
process (pstm_hip_set_tuning, pstm_hip_set_pipeline).
"""

from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass

from pm_amd.kernels import _abi

DEFAULT_CANDIDATES = [(mw, b, pipe) for mw in (1, 2, 4) for b in (4, 8, 16) for pipe in (0, 1)]


@dataclass
class TuneResult:
    mw: int
    batch: int
    pipe: int
    ms: float
    table: list

    def as_dict(self) -> dict:
        return {"mw": self.mw, "batch": self.batch, "pipe": self.pipe, "ms": round(self.ms, 3),
                "table": [dict(r) for r in self.table]}


def _time_plan(lib, ex, tiles) -> tuple[float, dict]:
    plan = ex.plan(tiles=tiles)
    try:
        lib.pstm_hip_profile_reset(1)
        plan.run(use_graph=False)
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        lib.pstm_hip_profile_read(ctypes.byref(ms), ctypes.byref(n))
        lib.pstm_hip_profile_reset(0)
        return ms.value / max(n.value, 1), plan.info()
    finally:
        plan.destroy()


def autotune(kernel, ex, candidates=None, n_tiles: int = 4, verbose: bool = True) -> TuneResult:
    """Time every candidate (mw, batch, pipe) as one plan over the job's
    n_tiles busiest tiles (a single tile's launch has too few work items per
    workgroup to show the tail-free shape a full plan runs with) and keep the
    fastest.  The launcher may clamp a candidate to a smaller batch when the
    rows do not fit in LDS or in the staging registers; the table shows the
    shape it ran.  Image and fold accumulate into the executor's buffers:
    reset them afterwards."""
    lib = _abi.load_library()
    order = sorted(range(len(ex.tiles)), key=lambda i: -int(ex.tiles[i].rows.shape[0]))[:n_tiles]
    table = []
    lib.pstm_hip_set_tuning(0, 0)
    lib.pstm_hip_set_pipeline(-1)
    _time_plan(lib, ex, order[:1])  # warm-up (code load)
    for cand in candidates or DEFAULT_CANDIDATES:
        mw, batch, pipe = (tuple(cand) + (-1,))[:3]
        if lib.pstm_hip_set_tuning(mw, batch) != 0 or lib.pstm_hip_set_pipeline(pipe) != 0:
            continue
        ms, shape = _time_plan(lib, ex, order)
        row = {"mw_cap": mw, "batch_cap": batch, "pipe": pipe, "mw": shape["mw"], "batch": shape["batch"],
               "items": shape["items"], "ms": round(ms, 3)}
        table.append(row)
        if verbose:
            print(f"[autotune] {row}", file=sys.stderr, flush=True)
    best = min(table, key=lambda r: r["ms"])
    lib.pstm_hip_set_tuning(best["mw"], best["batch_cap"])
    lib.pstm_hip_set_pipeline(best["pipe"])
    ex.reset()
    return TuneResult(best["mw"], best["batch"], best["pipe"], best["ms"], table)


@dataclass
class TileTuneResult:
    """Best tile edge (output bins per tile side) and launch shape for a
    fixed output region, with every measured (edge, shape) row."""

    edge: int
    mw: int
    batch: int
    pipe: int
    ms: float
    table: list

    def as_dict(self) -> dict:
        return {"edge": self.edge, "mw": self.mw, "batch": self.batch, "pipe": self.pipe, "ms": round(self.ms, 3),
                "table": [dict(r) for r in self.table]}


def region_tiles(ex, x0: int, x1: int, y0: int, y1: int) -> list[int]:
    """Indices of the executor's tiles inside the pillar region [x0, x1) x [y0, y1)."""
    return [i for i, t in enumerate(ex.tiles) if t.x0 >= x0 and t.x1 <= x1 and t.y0 >= y0 and t.y1 <= y1]


def autotune_tiles(kernel, job, edges=(16, 32, 64), region: int = 64, candidates=None, config=None,
                   max_edge: int | None = None, verbose: bool = True) -> TileTuneResult:
    """The output-bin dimension of the tile shape: for each tile edge the job
    is split into tiles of that edge, the tiles covering the same central
    region of region x region pillars are timed as one plan for every
    (mw, batch, pipe) candidate, and the fastest (edge, shape) wins.  The
    time covers the same pillars, traces and samples for every edge, so the
    rows compare directly: a small edge gives tighter trace lists (less halo
    outside the aperture of most pillars), a large one fewer, longer lists.
    The reference picks tile sizes from memory alone
    (pstm/pipeline/tile_planner.py:26-88).  The winner is the fastest row
    with edge <= max_edge (all rows by default); its launch shape is set
    (pstm_hip_set_tuning / _set_pipeline) and the caller builds its executor
    with its edge."""
    from dataclasses import replace

    from pm_amd.pipeline.executor import TileExecutor

    lib = _abi.load_library()
    nx, ny = len(job.x_axis), len(job.y_axis)
    if region > min(nx, ny) or any(region % e for e in edges):
        raise ValueError(f"region {region} must fit the grid and be a multiple of every edge {edges}")
    x0, y0 = (nx - region) // 2, (ny - region) // 2
    x0 -= x0 % max(edges)
    y0 -= y0 % max(edges)
    table = []
    for edge in edges:
        ex = TileExecutor(kernel, replace(job, tile=edge), config=config)
        tiles = region_tiles(ex, x0, x0 + region, y0, y0 + region)
        lib.pstm_hip_set_tuning(0, 0)
        lib.pstm_hip_set_pipeline(-1)
        _time_plan(lib, ex, tiles[:1])     # warm-up
        for cand in candidates or [(3, 16, 0), (4, 16, 0)]:
            mw, batch, pipe = (tuple(cand) + (-1,))[:3]
            if lib.pstm_hip_set_tuning(mw, batch) != 0 or lib.pstm_hip_set_pipeline(pipe) != 0:
                continue
            ms, shape = _time_plan(lib, ex, tiles)
            row = {"edge": edge, "tiles": len(tiles), "mw_cap": mw, "batch_cap": batch, "pipe": pipe,
                   "mw": shape["mw"], "batch": shape["batch"], "items": shape["items"], "ms": round(ms, 3)}
            table.append(row)
            if verbose:
                print(f"[autotune] {row}", file=sys.stderr, flush=True)
        del ex
    best = min((r for r in table if max_edge is None or r["edge"] <= max_edge), key=lambda r: r["ms"])
    lib.pstm_hip_set_tuning(best["mw"], best["batch_cap"])
    lib.pstm_hip_set_pipeline(best["pipe"])
    return TileTuneResult(best["edge"], best["mw"], best["batch"], best["pipe"], best["ms"], table)
