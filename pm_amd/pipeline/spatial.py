"""Spatial trace selection for output tiles.

The reference builds a cKDTree over midpoints and, per tile, returns the
traces whose midpoint lies in the tile rectangle padded by the aperture
(pstm/data/spatial_index.py:178-230, 423-460).  Here traces are kept in
Z-order (Morton) of their midpoints, so that the rows of one trace batch are
spatial neighbours (whole batches fall outside a workgroup's aperture and are
skipped), and the tile query is a vectorised rectangle test that returns rows
in that order.
"""

from __future__ import annotations

import numpy as np


def _spread_bits(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & np.uint64(0xFFFF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
    v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
    v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
    v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
    return v


def morton_order(mx: np.ndarray, my: np.ndarray, cell: float = 25.0) -> np.ndarray:
    """Permutation sorting traces by the Z-order key of their midpoint cell
    (stable, so equal cells keep index order)."""
    mx = np.asarray(mx, dtype=np.float64)
    my = np.asarray(my, dtype=np.float64)
    if len(mx) == 0:
        return np.zeros(0, dtype=np.int64)
    ix = np.floor((mx - mx.min()) / cell).astype(np.int64).clip(0, 0xFFFF)
    iy = np.floor((my - my.min()) / cell).astype(np.int64).clip(0, 0xFFFF)
    key = _spread_bits(ix) | (_spread_bits(iy) << np.uint64(1))
    return np.argsort(key, kind="stable")


def query_traces_for_tile(mx, my, x_min: float, x_max: float, y_min: float, y_max: float,
                          aperture_radius: float):
    """Rows whose midpoint lies in [x_min - r, x_max + r] x [y_min - r,
    y_max + r] (closed), ascending.  numpy arrays give a numpy int64 array,
    torch tensors a torch int32 tensor on the same device."""
    r = aperture_radius
    m = (mx >= x_min - r) & (mx <= x_max + r) & (my >= y_min - r) & (my <= y_max + r)
    if isinstance(m, np.ndarray):
        return np.nonzero(m)[0]
    import torch

    return torch.nonzero(m).flatten().to(torch.int32)
