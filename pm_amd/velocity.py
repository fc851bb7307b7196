"""3-D RMS velocity cube resampled onto the output grid.

The reference builds a SciPy ``RegularGridInterpolator(method="linear")`` over
the cube with coordinates clamped to the cube's extent
(pstm/data/velocity_model.py:375-396, 559-679) and hands each tile a
(tile_nx, tile_ny, nt) float64 slice of the precomputed volume
(velocity_model.py:871-950).  Here the same trilinear, clamped resampling runs
as tensor arithmetic on the device that holds the output grid.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class VelocityCube:
    """v(x, y, t) samples on a regular grid: data (nx, ny, nt) float64."""

    data: np.ndarray
    x_axis: np.ndarray
    y_axis: np.ndarray
    t_axis_ms: np.ndarray

    def resample(self, x, y, t_ms, device="cpu"):
        """Trilinear interpolation at the grid x (nx,) x y (ny,) x t (nt,)
        with clamping, returned as a (nx, ny, nt) float64 torch tensor."""
        import torch

        cube = torch.as_tensor(np.ascontiguousarray(self.data), dtype=torch.float64, device=device)

        def weights(axis, pts):
            axis = np.asarray(axis, dtype=np.float64)
            p = np.clip(np.asarray(pts, dtype=np.float64), axis[0], axis[-1])
            i = np.clip(np.searchsorted(axis, p, side="right") - 1, 0, len(axis) - 2)
            f = (p - axis[i]) / (axis[i + 1] - axis[i])
            return (torch.as_tensor(i, device=device), torch.as_tensor(f, dtype=torch.float64, device=device))

        ix, fx = weights(self.x_axis, x)
        iy, fy = weights(self.y_axis, y)
        it, ft = weights(self.t_axis_ms, t_ms)
        out = torch.zeros((len(ix), len(iy), len(it)), dtype=torch.float64, device=device)
        for dx in (0, 1):
            wx = (fx if dx else 1 - fx)[:, None, None]
            cx = cube[ix + dx]
            for dy in (0, 1):
                wy = (fy if dy else 1 - fy)[None, :, None]
                cxy = cx[:, iy + dy]
                for dz in (0, 1):
                    wz = (ft if dz else 1 - ft)[None, None, :]
                    out += wx * wy * wz * cxy[:, :, it + dz]
        return out

    def at(self, x, y, t_ms) -> np.ndarray:
        """Trilinear values at scattered points (numpy, float64)."""
        x = np.atleast_1d(np.asarray(x, float))
        y = np.atleast_1d(np.asarray(y, float))
        t = np.atleast_1d(np.asarray(t_ms, float))
        res = np.empty(len(x))
        for n in range(len(x)):
            res[n] = float(self.resample([x[n]], [y[n]], [t[n]])[0, 0, 0])
        return res
