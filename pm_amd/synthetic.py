"""Seeded synthetic migration jobs: the workloads the backend is measured on.

Geometry and event parameters come from ``numpy.random.default_rng(seed)``
(PCG64) in float64; traces are synthesised as tensor arithmetic on the chosen
device (the 200k x 1501 job is built on the GPU in well under a second) and
cast to float32.  Ricker wavelet ``(1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2)``
as the reference's synthetic generator uses (tests/fixtures/synthetic.py:21-44).

Jobs:
  * ``point_diffractor_job`` — one diffractor at (1000, 1000) m, t0 0.8 s,
    v 2000 m/s, 4096 traces x 1001 samples, 64 x 64 x 500 output at 25 m,
    aperture 1000 m, weights on, no anti-aliasing.
  * ``reflector_job`` — 200,000 traces x 1501 samples, 40 dipping planes and
    200 diffractors, 3-D v_rms cube 64 x 64 x 151 with +-10 % lateral bumps,
    256 x 256 x 1000 output at 12.5 m in 64 x 64 tiles, aperture 2500 m, dip
    60 deg, triangle anti-aliasing.
Sizes are parameters so that tests can run the same recipes scaled down.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

from pm_amd.kernels.base import KernelConfig
from pm_amd.velocity import VelocityCube


@dataclass
class SyntheticJob:
    """Everything one migration job needs, amplitudes as a torch tensor."""

    name: str
    amplitudes: Any                 # torch (n_traces, n_samples) f32
    source_x: np.ndarray
    source_y: np.ndarray
    receiver_x: np.ndarray
    receiver_y: np.ndarray
    dt_ms: float
    t_start_ms: float
    x_axis: np.ndarray
    y_axis: np.ndarray
    t_axis_ms: np.ndarray
    config: KernelConfig
    vrms_1d: np.ndarray | None = None        # (nt,) when the velocity is 1-D
    cube: VelocityCube | None = None         # 3-D v_rms model
    tile: int = 0                            # output tile edge (0: one tile)
    meta: dict = field(default_factory=dict)

    @property
    def midpoint_x(self) -> np.ndarray:
        return (self.source_x + self.receiver_x) / 2.0

    @property
    def midpoint_y(self) -> np.ndarray:
        return (self.source_y + self.receiver_y) / 2.0

    @property
    def n_traces(self) -> int:
        return int(self.amplitudes.shape[0])

    @property
    def n_samples(self) -> int:
        return int(self.amplitudes.shape[1])

    def tiles(self) -> list[tuple[int, int, int, int]]:
        """(x0, x1, y0, y1) index ranges of the output tiles."""
        nx, ny = len(self.x_axis), len(self.y_axis)
        t = self.tile or max(nx, ny)
        return [(x0, min(x0 + t, nx), y0, min(y0 + t, ny))
                for x0 in range(0, nx, t) for y0 in range(0, ny, t)]


def ricker(tau, f_hz):
    a = (np.pi * f_hz * tau) ** 2
    if isinstance(a, np.ndarray) or np.isscalar(a):
        return (1.0 - 2.0 * a) * np.exp(-a)
    return (1.0 - 2.0 * a) * a.neg().exp()


def add_events(traces, t_event, amp, f_hz: float, dt_s: float, t_start_s: float = 0.0):
    """traces[j, n] += amp[j] * ricker(t_start + n dt - t_event[j]) over the
    samples where the wavelet is above 1e-7 of its peak."""
    import torch

    n_tr, ns = traces.shape
    half = int(np.ceil(4.5 / (np.pi * f_hz * dt_s)))
    rows = torch.arange(n_tr, device=traces.device)
    centre = torch.round((t_event - t_start_s) / dt_s).long()
    for w in range(-half, half + 1):
        idx = centre + w
        ok = (idx >= 0) & (idx < ns)
        tau = t_start_s + idx.double() * dt_s - t_event
        val = torch.where(ok, amp * ricker(tau, f_hz), torch.zeros_like(tau))
        traces[rows, idx.clamp(0, ns - 1)] += val.to(traces.dtype)
    return traces


def _geometry(rng, n, lo, hi, max_offset):
    mx = rng.uniform(lo, hi, n)
    my = rng.uniform(lo, hi, n)
    h = rng.uniform(0.0, max_offset, n)
    az = rng.uniform(0.0, 2.0 * np.pi, n)
    c, s = np.cos(az), np.sin(az)
    return mx - 0.5 * h * c, my - 0.5 * h * s, mx + 0.5 * h * c, my + 0.5 * h * s


def point_diffractor_job(seed: int = 1, n_traces: int = 4096, n_samples: int = 1001,
                         n_out: int = 64, nt: int = 500, device: str = "cpu") -> SyntheticJob:
    import torch

    rng = np.random.default_rng(seed)
    sx, sy, rx, ry = _geometry(rng, n_traces, 0.0, 2000.0, 1500.0)
    dx_, dy_, t0d, v = 1000.0, 1000.0, 0.8, 2000.0
    dt_s = 0.002
    t_d = (np.sqrt((t0d / 2) ** 2 + ((sx - dx_) ** 2 + (sy - dy_) ** 2) / v ** 2)
           + np.sqrt((t0d / 2) ** 2 + ((rx - dx_) ** 2 + (ry - dy_) ** 2) / v ** 2))
    tr = torch.zeros((n_traces, n_samples), dtype=torch.float64, device=device)
    add_events(tr, torch.as_tensor(t_d, device=device), torch.ones(n_traces, dtype=torch.float64,
               device=device), 25.0, dt_s)
    peak = float(tr.abs().max())
    noise = torch.as_tensor(rng.standard_normal((n_traces, n_samples)), device=device)
    tr += 0.01 * peak * noise
    x = 200.0 + 25.0 * np.arange(n_out)
    cfg = KernelConfig(max_aperture_m=1000.0, min_aperture_m=1000.0, max_dip_degrees=45.0,
                       taper_fraction=0.1, apply_spreading=True, apply_obliquity=True,
                       grid_dx=25.0, grid_dy=25.0)
    return SyntheticJob("point_diffractor", tr.float().contiguous(), sx, sy, rx, ry, 2.0, 0.0, x,
                        x.copy(), 2.0 * np.arange(nt), cfg, vrms_1d=np.full(nt, v),
                        meta={"diffractor": (dx_, dy_, t0d)})


def lateral_cube(x_lo: float = -500.0, x_hi: float = 3687.5, n_xy: int = 64, n_t: int = 151,
                 dt_ms: float = 20.0) -> VelocityCube:
    """v = (1500 + 600 t0)(1 + 0.1 (g1 - g2)), two Gaussian bumps of sigma
    800 m centred at (1000, 1000) and (2500, 2200) m."""
    x = np.linspace(x_lo, x_hi, n_xy)
    t = dt_ms * np.arange(n_t)
    X, Y = np.meshgrid(x, x, indexing="ij")
    sig = 800.0
    g1 = np.exp(-((X - 1000.0) ** 2 + (Y - 1000.0) ** 2) / (2 * sig ** 2))
    g2 = np.exp(-((X - 2500.0) ** 2 + (Y - 2200.0) ** 2) / (2 * sig ** 2))
    v = (1500.0 + 600.0 * t / 1000.0)[None, None, :] * (1.0 + 0.1 * (g1 - g2))[:, :, None]
    return VelocityCube(v, x, x.copy(), t)


def reflector_job(seed: int = 2, n_traces: int = 200_000, n_samples: int = 1501,
                  n_out: int = 256, nt: int = 1000, d_out: float = 12.5, tile: int = 64,
                  n_planes: int = 40, n_diffractors: int = 200, antialias: str = "triangle",
                  device: str = "cpu") -> SyntheticJob:
    rng = np.random.default_rng(seed)
    x_hi = d_out * (n_out - 1)
    lo, hi = -500.0, x_hi + 500.0
    sx, sy, rx, ry = _geometry(rng, n_traces, lo, hi, 3000.0)
    cube = lateral_cube(lo, hi)
    tr, meta = _reflector_traces(rng, (sx, sy, rx, ry), n_samples, cube, x_hi, n_planes, n_diffractors, device)
    x = d_out * np.arange(n_out)
    cfg = KernelConfig(max_aperture_m=2500.0, min_aperture_m=500.0, max_dip_degrees=60.0,
                       taper_fraction=0.1, apply_spreading=True, apply_obliquity=True,
                       grid_dx=d_out, grid_dy=d_out, antialias=antialias)
    return SyntheticJob("reflectors", tr, sx, sy, rx, ry, 2.0, 0.0, x,
                        x.copy(), 2.0 * np.arange(nt), cfg, cube=cube, tile=tile, meta=meta)


def _reflector_traces(rng, geom, n_samples: int, cube, x_hi: float, n_planes: int, n_diffractors: int,
                      device: str):
    """Dipping planes (mirror-image time in the plane-centre velocity) and
    point diffractors (DSR time in the local velocity), 30 Hz Ricker, f32."""
    import torch

    dt_s = 0.002
    f = 30.0
    tr = torch.zeros((len(geom[0]), n_samples), dtype=torch.float64, device=device)
    S = [torch.as_tensor(a, device=device) for a in geom]
    xc = yc = 0.5 * x_hi
    dips = rng.uniform(0.0, np.deg2rad(30.0), n_planes)
    strikes = rng.uniform(0.0, 2.0 * np.pi, n_planes)
    t0c = rng.uniform(0.3, 2.8, n_planes)
    amps = rng.uniform(0.3, 1.0, n_planes) * rng.choice([-1.0, 1.0], n_planes)
    vc = cube.at(np.full(n_planes, xc), np.full(n_planes, yc), 1000.0 * t0c)
    for i in range(n_planes):
        n_vec = np.array([np.sin(dips[i]) * np.cos(strikes[i]),
                          np.sin(dips[i]) * np.sin(strikes[i]), np.cos(dips[i])])
        p0 = np.array([xc, yc, vc[i] * t0c[i] / 2.0])  # plane point under the centre
        d_s = (S[0] - p0[0]) * n_vec[0] + (S[1] - p0[1]) * n_vec[1] + (0.0 - p0[2]) * n_vec[2]
        mx_ = S[0] - 2 * d_s * n_vec[0]
        my_ = S[1] - 2 * d_s * n_vec[1]
        mz_ = -2 * d_s * n_vec[2]
        t = torch.sqrt((mx_ - S[2]) ** 2 + (my_ - S[3]) ** 2 + mz_ ** 2) / vc[i]
        add_events(tr, t, torch.full_like(t, amps[i]), f, dt_s)
    xd = rng.uniform(0.0, x_hi, n_diffractors)
    yd = rng.uniform(0.0, x_hi, n_diffractors)
    t0d = rng.uniform(0.2, 2.8, n_diffractors)
    ad = rng.uniform(0.5, 1.0, n_diffractors)
    vd = cube.at(xd, yd, 1000.0 * t0d)
    for i in range(n_diffractors):
        q = (t0d[i] / 2) ** 2
        t = (torch.sqrt(q + ((S[0] - xd[i]) ** 2 + (S[1] - yd[i]) ** 2) / vd[i] ** 2)
             + torch.sqrt(q + ((S[2] - xd[i]) ** 2 + (S[3] - yd[i]) ** 2) / vd[i] ** 2))
        add_events(tr, t, torch.full_like(t, ad[i]), f, dt_s)
    return tr.float().contiguous(), {"diffractors": np.stack([xd, yd, t0d], 1)}


def bin_counts(rng, n_bins: int = 40, total: int = 1_600_000, median: int = 30_000, lo: int = 5_000,
               hi: int = 120_000) -> np.ndarray:
    """Lognormal trace counts per common-offset bin, clip(median e^{0.6 z},
    lo, hi), rescaled to sum to about `total`."""
    c = np.clip(median * np.exp(0.6 * rng.standard_normal(n_bins)), lo, hi)
    return np.maximum(1, np.round(c * total / c.sum())).astype(np.int64)


def bins_job(seed: int = 3, n_bins: int = 40, total: int = 1_600_000, n_samples: int = 1501,
             n_out: int = 256, nt: int = 1000, d_out: float = 12.5, tile: int = 64, bin_width: float = 50.0,
             n_planes: int = 40, n_diffractors: int = 200, device: str = "cpu") -> SyntheticJob:
    """Common-offset bins b = 0 .. n_bins-1 (offsets in [50 b, 50 b + 50) m,
    uniform midpoints and azimuths) with lognormal counts, the reflector
    job's reflectivity and velocity, plain migration (no anti-aliasing) into
    one shared image.  Traces are stored bin by bin; meta["bin_start"] holds
    the first row of each bin."""
    rng = np.random.default_rng(seed)
    counts = bin_counts(rng, n_bins, total, median=max(1, int(30_000 * total / 1_600_000)),
                        lo=max(1, int(5_000 * total / 1_600_000)), hi=max(1, int(120_000 * total / 1_600_000)))
    x_hi = d_out * (n_out - 1)
    lo, hi = -500.0, x_hi + 500.0
    parts = []
    for b, n in enumerate(counts):
        mx, my = rng.uniform(lo, hi, n), rng.uniform(lo, hi, n)
        h = rng.uniform(bin_width * b, bin_width * (b + 1), n)
        az = rng.uniform(0.0, 2.0 * np.pi, n)
        parts.append((mx - 0.5 * h * np.cos(az), my - 0.5 * h * np.sin(az),
                      mx + 0.5 * h * np.cos(az), my + 0.5 * h * np.sin(az)))
    sx, sy, rx, ry = (np.concatenate([p[i] for p in parts]) for i in range(4))
    cube = lateral_cube(lo, hi)
    tr, meta = _reflector_traces(rng, (sx, sy, rx, ry), n_samples, cube, x_hi, n_planes, n_diffractors, device)
    meta["bin_start"] = np.concatenate([[0], np.cumsum(counts)])
    x = d_out * np.arange(n_out)
    cfg = KernelConfig(max_aperture_m=2500.0, min_aperture_m=500.0, max_dip_degrees=60.0,
                       taper_fraction=0.1, apply_spreading=True, apply_obliquity=True,
                       grid_dx=d_out, grid_dy=d_out, antialias="none")
    return SyntheticJob("bins", tr, sx, sy, rx, ry, 2.0, 0.0, x, x.copy(), 2.0 * np.arange(nt), cfg,
                        cube=cube, tile=tile, meta=meta)


def vti_model(n_xy: int = 128, n_t: int = 201, x_lo: float = -500.0, x_hi: float = 6887.5,
              t_hi_ms: float = 4000.0) -> tuple[VelocityCube, VelocityCube]:
    """Smooth 3-D v_rms (1500-4500 m/s) and eta (0-0.2) fields on one grid:
    v = 1500 + 900 t0 (1 + 0.15 sin(2 pi x / Lx) cos(2 pi y / Ly)) and
    eta = 0.1 + 0.08 sin(2 pi (x + y) / L) * min(1, t0), clipped."""
    x = np.linspace(x_lo, x_hi, n_xy)
    t = np.linspace(0.0, t_hi_ms, n_t)
    X, Y, T = np.meshgrid(x, x, t / 1000.0, indexing="ij")
    span = x_hi - x_lo
    v = 1500.0 + 900.0 * T * (1.0 + 0.15 * np.sin(2 * np.pi * X / span) * np.cos(2 * np.pi * Y / span))
    eta = 0.1 + 0.08 * np.sin(2 * np.pi * (X + Y) / (1.3 * span)) * np.minimum(1.0, T)
    return (VelocityCube(np.clip(v, 1500.0, 4500.0), x, x.copy(), t),
            VelocityCube(np.clip(eta, 0.0, 0.2), x, x.copy(), t))


def _vti_leg(tau, h, v, eta):
    """One-way VTI time (anisotropy_vti.py:181-220) as f64 tensor arithmetic,
    for synthesising data (tau > 0, v > 0)."""
    h2, v2, t02 = h * h, v * v, tau * tau
    den = v2 * (t02 * v2 + (1.0 + 2.0 * eta) * h2)
    return (t02 + h2 / v2 - 2.0 * eta * h2 * h2 / den).clamp_min(0.0).sqrt()


def vti_diffractor_job(seed: int = 4, n_traces: int = 500_000, n_samples: int = 2001, n_out: int = 512,
                       nt: int = 1500, d_out: float = 12.5, tile: int = 64, n_diffractors: int = 300,
                       device: str = "cpu") -> SyntheticJob:
    """VTI job: random diffractors at the exact two-leg VTI time with the local
    v_rms and eta, 30 Hz Ricker; output n_out x n_out bins x nt samples."""
    import torch

    rng = np.random.default_rng(seed)
    x_hi = d_out * (n_out - 1)
    lo, hi = -500.0, x_hi + 500.0
    sx, sy, rx, ry = _geometry(rng, n_traces, lo, hi, 3000.0)
    vc, ec = vti_model(x_lo=lo, x_hi=hi)
    dt_s = 0.002
    tr = torch.zeros((n_traces, n_samples), dtype=torch.float64, device=device)
    S = [torch.as_tensor(a, device=device) for a in (sx, sy, rx, ry)]
    xd = rng.uniform(0.0, x_hi, n_diffractors)
    yd = rng.uniform(0.0, x_hi, n_diffractors)
    t0d = rng.uniform(0.2, 0.98 * nt * dt_s, n_diffractors)
    ad = rng.uniform(0.5, 1.0, n_diffractors)
    vd = vc.at(xd, yd, 1000.0 * t0d)
    ed = ec.at(xd, yd, 1000.0 * t0d)
    for i in range(n_diffractors):
        tau = torch.full_like(S[0], t0d[i] / 2)
        hs = ((S[0] - xd[i]) ** 2 + (S[1] - yd[i]) ** 2).sqrt()
        hr = ((S[2] - xd[i]) ** 2 + (S[3] - yd[i]) ** 2).sqrt()
        vv = torch.full_like(tau, vd[i])
        ee = torch.full_like(tau, ed[i])
        t = _vti_leg(tau, hs, vv, ee) + _vti_leg(tau, hr, vv, ee)
        add_events(tr, t, torch.full_like(t, ad[i]), 30.0, dt_s)
    x = d_out * np.arange(n_out)
    cfg = KernelConfig(max_aperture_m=2500.0, min_aperture_m=500.0, max_dip_degrees=60.0,
                       taper_fraction=0.1, apply_spreading=True, apply_obliquity=True,
                       grid_dx=d_out, grid_dy=d_out, kernel_type="anisotropic_vti", vti_enabled=True)
    job = SyntheticJob("vti_diffractors", tr.float().contiguous(), sx, sy, rx, ry, 2.0, 0.0, x, x.copy(),
                       2.0 * np.arange(nt), cfg, cube=vc, tile=tile,
                       meta={"eta_cube": ec, "diffractors": np.stack([xd, yd, t0d], 1)})
    return job


def vti_bins_job(seed: int = 7, n_bins: int = 60, total: int = 2_400_000, n_samples: int = 2001,
                 n_out: int = 1024, nt: int = 2000, d_out: float = 12.5, tile: int = 16, bin_width: float = 50.0,
                 n_diffractors: int = 300, device: str = "cpu") -> SyntheticJob:
    """The largest single-GPU job: common-offset bins b = 0 .. n_bins-1
    (offsets in [50 b, 50 b + 50) m, uniform midpoints and azimuths) with
    lognormal counts (about `total` traces, stored bin by bin,
    meta["bin_start"]), the VTI model of vti_diffractor_job over the survey,
    random diffractors at their exact two-leg VTI times (30 Hz Ricker,
    amplitudes accumulated in f32); output n_out x n_out bins x nt samples
    at 2 ms."""
    import torch

    rng = np.random.default_rng(seed)
    scale = total / 1_600_000
    counts = bin_counts(rng, n_bins, total, median=max(1, int(30_000 * scale)), lo=max(1, int(5_000 * scale)),
                        hi=max(1, int(120_000 * scale)))
    x_hi = d_out * (n_out - 1)
    lo, hi = -500.0, x_hi + 500.0
    parts = []
    for b, n in enumerate(counts):
        mx, my = rng.uniform(lo, hi, n), rng.uniform(lo, hi, n)
        h = rng.uniform(bin_width * b, bin_width * (b + 1), n)
        az = rng.uniform(0.0, 2.0 * np.pi, n)
        parts.append((mx - 0.5 * h * np.cos(az), my - 0.5 * h * np.sin(az),
                      mx + 0.5 * h * np.cos(az), my + 0.5 * h * np.sin(az)))
    sx, sy, rx, ry = (np.concatenate([p[i] for p in parts]) for i in range(4))
    vc, ec = vti_model(x_lo=lo, x_hi=hi)
    dt_s = 0.002
    tr = torch.zeros((len(sx), n_samples), dtype=torch.float32, device=device)
    S = [torch.as_tensor(a, device=device) for a in (sx, sy, rx, ry)]
    xd = rng.uniform(0.0, x_hi, n_diffractors)
    yd = rng.uniform(0.0, x_hi, n_diffractors)
    t0d = rng.uniform(0.2, 0.98 * nt * dt_s, n_diffractors)
    ad = rng.uniform(0.5, 1.0, n_diffractors)
    vd = vc.at(xd, yd, 1000.0 * t0d)
    ed = ec.at(xd, yd, 1000.0 * t0d)
    for i in range(n_diffractors):
        tau = torch.full_like(S[0], t0d[i] / 2)
        hs = ((S[0] - xd[i]) ** 2 + (S[1] - yd[i]) ** 2).sqrt()
        hr = ((S[2] - xd[i]) ** 2 + (S[3] - yd[i]) ** 2).sqrt()
        vv = torch.full_like(tau, vd[i])
        ee = torch.full_like(tau, ed[i])
        t = _vti_leg(tau, hs, vv, ee) + _vti_leg(tau, hr, vv, ee)
        add_events(tr, t, torch.full_like(t, ad[i]), 30.0, dt_s)
    x = d_out * np.arange(n_out)
    cfg = KernelConfig(max_aperture_m=2500.0, min_aperture_m=500.0, max_dip_degrees=60.0,
                       taper_fraction=0.1, apply_spreading=True, apply_obliquity=True,
                       grid_dx=d_out, grid_dy=d_out, kernel_type="anisotropic_vti", vti_enabled=True)
    return SyntheticJob("vti_bins", tr, sx, sy, rx, ry, 2.0, 0.0, x, x.copy(), 2.0 * np.arange(nt), cfg,
                        cube=vc, tile=tile,
                        meta={"eta_cube": ec, "bin_start": np.concatenate([[0], np.cumsum(counts)]),
                              "diffractors": np.stack([xd, yd, t0d], 1)})


def curved_diffractor_job(seed: int = 6, n_traces: int = 200_000, n_samples: int = 1501, n_out: int = 128,
                          nt: int = 1000, d_out: float = 12.5, tile: int = 32, n_diffractors: int = 120,
                          v0: float = 1600.0, k: float = 0.6, device: str = "cpu") -> SyntheticJob:
    """Curved-ray job: random diffractors at the reference's curved-ray DSR
    time (pstm/algorithm/curved_ray.py:287-316: one-way legs of
    curved_ray_traveltime at the depth z(t0 / 2) of V(z) = v0 + k z), 30 Hz
    Ricker; output n_out x n_out bins x nt samples; v_rms for the aperture is
    the velocity at the image depth, v0 e^{k t0 / 2}."""
    import torch

    from pm_amd.surrogate import curved_leg_times

    rng = np.random.default_rng(seed)
    x_hi = d_out * (n_out - 1)
    lo, hi = -500.0, x_hi + 500.0
    sx, sy, rx, ry = _geometry(rng, n_traces, lo, hi, 3000.0)
    dt_s = 0.002
    tr = torch.zeros((n_traces, n_samples), dtype=torch.float64, device=device)
    S = [torch.as_tensor(a, device=device) for a in (sx, sy, rx, ry)]
    xd = rng.uniform(0.0, x_hi, n_diffractors)
    yd = rng.uniform(0.0, x_hi, n_diffractors)
    t0d = rng.uniform(0.2, 0.98 * nt * dt_s, n_diffractors)
    ad = rng.uniform(0.5, 1.0, n_diffractors)
    for i in range(n_diffractors):
        tau = torch.full_like(S[0], t0d[i] / 2)
        hs = ((S[0] - xd[i]) ** 2 + (S[1] - yd[i]) ** 2).sqrt()
        hr = ((S[2] - xd[i]) ** 2 + (S[3] - yd[i]) ** 2).sqrt()
        t = curved_leg_times(tau, hs, v0, k) + curved_leg_times(tau, hr, v0, k)
        add_events(tr, t, torch.full_like(t, ad[i]), 30.0, dt_s)
    x = d_out * np.arange(n_out)
    t_ms = 2.0 * np.arange(nt)
    cfg = KernelConfig(max_aperture_m=2500.0, min_aperture_m=500.0, max_dip_degrees=60.0,
                       taper_fraction=0.1, apply_spreading=True, apply_obliquity=True,
                       grid_dx=d_out, grid_dy=d_out, kernel_type="curved_ray", curved_ray_enabled=True,
                       curved_ray_v0=v0, curved_ray_k=k)
    return SyntheticJob("curved_diffractors", tr.float().contiguous(), sx, sy, rx, ry, 2.0, 0.0, x, x.copy(),
                        t_ms, cfg, vrms_1d=v0 * np.exp(k * t_ms / 2000.0), tile=tile,
                        meta={"diffractors": np.stack([xd, yd, t0d], 1), "v0": v0, "k": k})
