"""Launch-shape autotuner for the gather kernel.

The reference sizes tiles with a memory heuristic (pstm/pipeline/
tile_planner.py:26-88) and never measures.  Here the shape of one launch —
output samples per workgroup (256 x MW, MW sample blocks of 64 per wave) and
traces per LDS batch — is chosen by timing candidate shapes on the job's
busiest tile with HIP events on the launch stream, and the winner is set for
every later launch of the process (pstm_hip_set_tuning).
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


def _time_launch(lib, kernel, ex, tile, image, fold) -> float:
    lib.pstm_hip_profile_reset(1)
    kernel.migrate_device(ex.traces, image, fold, tile.x_axis, tile.y_axis, ex.t_axis, tile.vrms,
                          ex.config, trace_index=tile.rows)
    ms = ctypes.c_double()
    n = ctypes.c_longlong()
    lib.pstm_hip_profile_read(ctypes.byref(ms), ctypes.byref(n))
    lib.pstm_hip_profile_reset(0)
    return ms.value / max(n.value, 1)


def autotune(kernel, ex, candidates=None, verbose: bool = True) -> TuneResult:
    """Time every candidate (mw, batch, pipe) on the tile with the most traces
    and keep the fastest (the launcher may clamp a candidate to a smaller batch
    when the rows do not fit in LDS or in the pipelined staging registers; the
    table shows the shape it ran)."""
    import torch

    lib = _abi.load_library()
    tile = max(ex.tiles, key=lambda t: int(t.rows.shape[0]))
    image = torch.zeros_like(tile.image)
    fold = torch.zeros_like(tile.fold)
    table = []
    lib.pstm_hip_set_tuning(0, 0)
    lib.pstm_hip_set_pipeline(-1)
    _time_launch(lib, kernel, ex, tile, image, fold)  # warm-up (code load, workspace)
    for cand in candidates or DEFAULT_CANDIDATES:
        mw, batch, pipe = (tuple(cand) + (-1,))[:3]
        if lib.pstm_hip_set_tuning(mw, batch) != 0 or lib.pstm_hip_set_pipeline(pipe) != 0:
            continue
        ms = _time_launch(lib, kernel, ex, tile, image, fold)
        shape = _abi.last_launch(lib)
        row = {"mw_cap": mw, "batch_cap": batch, "pipe": pipe, **shape, "ms": round(ms, 3)}
        table.append(row)
        if verbose:
            print(f"[autotune] {row}", file=sys.stderr, flush=True)
    best = min(table, key=lambda r: r["ms"])
    lib.pstm_hip_set_tuning(best["mw"], best["batch_cap"])
    lib.pstm_hip_set_pipeline(best["pipe"])
    return TuneResult(best["mw"], best["batch"], best["pipe"], best["ms"], table)
