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
                  spreading=True, obliquity=True, aa=None):
    """Returns (image (nx,ny,nt) f64, fold2d (nx,ny) i64). vrms (nt,) or (nx,ny,nt).

    aa = (grid_dx, grid_dy, max_half_width): triangle anti-aliasing, written
    without the oracle's integrated traces: the operator dip dt/dx, dt/dy by
    central finite differences of the DSR time at the output point, half
    width L = max(|dt/dx| dx, |dt/dy| dy) / dt clamped to [1, max_half_width],
    and the triangles of the integer half widths floor(L) and floor(L) + 1 as
    explicit convolutions of the zero-extended trace, blended with weight
    L - floor(L) and interpolated linearly between the samples around the
    position."""
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
        if aa is None:
            a = amp[rows, i0] * (1 - f) + amp[rows, np.minimum(i0 + 1, ns - 1)] * f
        else:
            a = _aa_amplitude(amp, rows, ox, oy, sx, sy, rx, ry, t0, v, dt_in_ms / 1000.0, i0, f, aa)
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


def _aa_amplitude(amp, rows, ox, oy, sx, sy, rx, ry, t0, v, dt_s, i0, f, aa):
    gdx, gdy, lmax = aa
    h = 1e-3

    def t_at(px, py):
        return (np.sqrt((t0 / 2) ** 2 + ((px - sx) ** 2 + (py - sy) ** 2) / v ** 2)
                + np.sqrt((t0 / 2) ** 2 + ((px - rx) ** 2 + (py - ry) ** 2) / v ** 2))

    gx = (t_at(ox + h, oy) - t_at(ox - h, oy)) / (2 * h)
    gy = (t_at(ox, oy + h) - t_at(ox, oy - h)) / (2 * h)
    L = np.maximum(np.abs(gx) * gdx, np.abs(gy) * gdy) / dt_s
    L = np.minimum(np.where(L > 1.0, L, 1.0), float(lmax))
    lf = np.floor(L).astype(np.int64)
    lam = L - lf
    ntr, ns = amp.shape
    pad = int(lmax) + 2
    ext = np.zeros((ntr, ns + 2 * pad))
    ext[:, pad:pad + ns] = amp

    def tri(n, lint):
        acc = np.zeros(n.shape)
        for m in range(-int(lmax) - 1, int(lmax) + 2):
            wgt = np.maximum(lint - abs(m), 0) / (lint * lint)
            acc += wgt * ext[rows, n + m + pad]
        return acc

    def blended(n):
        return (1 - lam) * tri(n, lf) + lam * tri(n, lf + 1)

    return blended(i0) * (1 - f) + blended(i0 + 1) * f
