"""TEST INFRASTRUCTURE: NumPy brute force of the plain straight-ray tile
migration, written independently of the C++ oracle (time samples outer,
vectorised over pillars x traces) for tiny inputs.

Follows the same reference lines as the oracle
(pstm/kernels/numba_cpu_optimized.py:93-233,358-401) but with a different loop
order and summation order, so a transposed index, a dropped term or a wrong
sign in either implementation shows up as a disagreement.
"""

from __future__ import annotations

import numpy as np


def migrate_plain(amp, sx, sy, rx, ry, mx, my, w, x_axis, y_axis, t_axis_ms, vrms,
                  dt_in_ms, t_start_in_ms, max_ap, min_ap, dip_deg, taper,
                  spreading=True, obliquity=True):
    """Returns (image (nx,ny,nt) f64, fold2d (nx,ny) i64). vrms (nt,) or (nx,ny,nt)."""
    amp = np.asarray(amp, dtype=np.float64)
    ntr, ns = amp.shape
    X, Y = np.meshgrid(np.asarray(x_axis, float), np.asarray(y_axis, float), indexing="ij")
    ox = X[:, :, None]
    oy = Y[:, :, None]
    dm = np.sqrt((ox - mx[None, None, :]) ** 2 + (oy - my[None, None, :]) ** 2)
    inside = dm <= max_ap
    fold = inside.sum(axis=2)
    ds2 = (ox - sx) ** 2 + (oy - sy) ** 2
    dr2 = (ox - rx) ** 2 + (oy - ry) ** 2
    nt = len(t_axis_ms)
    image = np.zeros(X.shape + (nt,))
    tan_dip = np.tan(np.deg2rad(dip_deg))
    v_all = np.asarray(vrms, float)
    rows = np.arange(ntr)[None, None, :]
    for k in range(nt):
        t0 = t_axis_ms[k] / 1000.0
        v = v_all[k] if v_all.ndim == 1 else v_all[:, :, k][:, :, None]
        ap = np.clip(v * t0 * 0.5 * tan_dip, min_ap, max_ap)
        t = np.sqrt((t0 / 2) ** 2 + ds2 / v ** 2) + np.sqrt((t0 / 2) ** 2 + dr2 / v ** 2)
        u = (t - t_start_in_ms / 1000.0) / (dt_in_ms / 1000.0)
        ok = inside & (dm <= ap) & (u >= 0) & (u < ns - 1)
        i0 = np.where(ok, np.floor(u), 0).astype(np.int64)
        f = np.where(ok, u - i0, 0.0)
        a = amp[rows, i0] * (1 - f) + amp[rows, np.minimum(i0 + 1, ns - 1)] * f
        wt = np.broadcast_to(w[None, None, :], ok.shape).copy()
        ts = ap * (1 - taper)
        in_taper = dm > ts
        with np.errstate(divide="ignore", invalid="ignore"):
            x = (dm - ts) / (ap - ts)
        wt = np.where(in_taper, wt * 0.5 * (1 + np.cos(np.pi * x)), wt)
        late = t > 0.001
        if spreading:
            wt = np.where(late, wt / (v * t), wt)
        if obliquity:
            wt = np.where(late, wt * t0 / t, wt)
        image[:, :, k] = np.where(ok, a * wt, 0.0).sum(axis=2)
    return image, fold


def triangle_filter(trace: np.ndarray, n: int, L: int) -> float:
    """Explicit triangle smoothing of a zero-extended trace at integer sample n
    with integer half width L: (1/L^2) sum_{|m|<L} (L-|m|) A[n+m]."""
    acc = 0.0
    for m in range(-L + 1, L):
        j = n + m
        if 0 <= j < len(trace):
            acc += (L - abs(m)) * float(trace[j])
    return acc / (L * L)
