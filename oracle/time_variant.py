"""TEST INFRASTRUCTURE: restatement of the reference's time-variant sampling
(pstm/algorithm/time_variant.py) for checking pm_amd.time_variant and the HIP
resampling kernel; never part of the product path.

  * ``windows``  — compute_time_windows (:201-286) with interpolate_fmax
    (:124-148, SciPy PchipInterpolator clamped to the table ends),
    nearest_power_of_2 (:151-162) and compute_downsample_factor (:165-198);
    windows returned as (t_start, t_end, dt, factor, first, last, f_max)
    tuples;
  * ``resample`` — resample_to_uniform (:354-397): np.interp of every pillar
    onto k * base_dt, k < nt, f64, left zero when there is one sample or none.
"""

from __future__ import annotations

import numpy as np
from scipy.interpolate import PchipInterpolator


def _fmax(times, freqs, t):
    if t <= times[0]:
        return freqs[0]
    if t >= times[-1]:
        return freqs[-1]
    return float(PchipInterpolator(times, freqs)(t))


def _pow2(n):
    if n <= 1:
        return 1
    p = 1
    while 2 * p <= n:
        p *= 2
    return 2 * p if abs(n - p) > abs(n - 2 * p) else p


def _factor(f, base_dt, lo, hi):
    return max(lo, min(hi, _pow2(int((1000.0 / (2.0 * f)) / base_dt))))


def windows(t_min, t_max, base_dt, times, freqs, lo=1, hi=8):
    out = []
    t, first = t_min, 0
    while t < t_max:
        f = _fmax(times, freqs, t)
        fac = _factor(f, base_dt, lo, hi)
        dte = base_dt * fac
        stop = t
        while stop < t_max:
            if _factor(_fmax(times, freqs, stop + dte), base_dt, lo, hi) != fac:
                break
            stop += dte
        if stop <= t:
            stop = min(t + dte, t_max)
        stop = min(stop, t_max)
        n = max(1, int((stop - t) / dte))
        stop = t + n * dte
        if stop > t_max:
            stop = t_max
            n = max(1, int((stop - t) / dte))
        out.append((t, stop, dte, fac, first, first + n, f))
        first += n
        t = stop
    return out


def tv_times(wins):
    return np.array([w[0] + i * w[2] for w in wins for i in range(w[5] - w[4])], dtype=np.float64)


def resample(tv_image, wins, base_dt, nt):
    img = np.asarray(tv_image, dtype=np.float64)
    xp = tv_times(wins)
    grid = np.arange(nt) * base_dt
    res = np.zeros(img.shape[:-1] + (nt,))
    if len(xp) > 1:
        for idx in np.ndindex(*img.shape[:-1]):
            res[idx] = np.interp(grid, xp, img[idx])
    return res
