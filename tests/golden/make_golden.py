"""Write golden input/output vectors from the reference's own PSTM kernel code.

Run in a container that has the reference checkout at /root/reference:

    python tests/golden/make_golden.py

The reference hot loop (pstm/kernels/numba_cpu_optimized.py:93-233, 358-401)
and the kernel the reference tests call (pstm/kernels/numba_cpu.py:219-366)
are decorated with numba's @njit.  numba is not installed here, so a stand-in
module whose ``njit`` returns the function unchanged and whose ``prange`` is
``range`` is put in ``sys.modules`` first; the reference functions then run as
plain CPython/NumPy FP64 code on tiny seeded inputs.  (Numba's ``fastmath``
re-association is lost, so the reference's own compiled results may differ
from these at the ~1e-15 relative level.)

The curved-ray and VTI scalar traveltimes (pstm/algorithm/curved_ray.py:
287-316, pstm/algorithm/anisotropy_vti.py:148-220) are pure NumPy and are
evaluated on a grid of inputs.

Only data (inputs and outputs) is written, to ``tests/golden/*.npz``; nothing
from the reference is copied into this repository.
"""

from __future__ import annotations

import sys
import types
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent


def _install_numba_standin() -> None:
    mod = types.ModuleType("numba")

    def njit(*args, **kwargs):
        if len(args) == 1 and callable(args[0]) and not kwargs:
            return args[0]
        return lambda f: f

    mod.njit = njit
    mod.jit = njit
    mod.prange = range
    sys.modules["numba"] = mod


def _geometry(rng, n, extent, max_offset):
    mx = rng.uniform(extent[0], extent[1], n)
    my = rng.uniform(extent[0], extent[1], n)
    h = rng.uniform(0.0, max_offset, n)
    az = rng.uniform(0.0, 2 * np.pi, n)
    sx = mx - 0.5 * h * np.cos(az)
    sy = my - 0.5 * h * np.sin(az)
    rx = mx + 0.5 * h * np.cos(az)
    ry = my + 0.5 * h * np.sin(az)
    return sx, sy, rx, ry


def _case(nco, nc, name, seed, *, n_traces=64, n_samples=101, dt_in=4.0,
          t_start_in=0.0, nx=8, ny=8, nt=50, dx=40.0, x0=0.0, dt_out=4.0,
          t0_out=0.0, vrms=None, extent=(-200.0, 480.0), max_offset=600.0,
          max_ap=400.0, min_ap=100.0, dip=45.0, taper=0.1, spreading=True,
          obliquity=True, weights=False, init_nonzero=False, snap=False):
    rng = np.random.default_rng(seed)
    sx, sy, rx, ry = _geometry(rng, n_traces, extent, max_offset)
    ox = x0 + dx * np.arange(nx, dtype=np.float64)
    oy = x0 + dx * np.arange(ny, dtype=np.float64)
    if snap:
        # put a few sources/receivers exactly on output pillars (t -> 0 edge)
        for j in range(min(4, n_traces)):
            sx[j] = ox[j % nx]
            sy[j] = oy[(2 * j) % ny]
            rx[j] = sx[j]
            ry[j] = sy[j]
    mx = (sx + rx) / 2.0
    my = (sy + ry) / 2.0
    amp = rng.standard_normal((n_traces, n_samples)).astype(np.float32)
    wtr = rng.uniform(0.5, 1.5, n_traces) if weights else np.ones(n_traces)
    ot = t0_out + dt_out * np.arange(nt, dtype=np.float64)
    if vrms is None:
        vrms = 1800.0 + 400.0 * ot / 1000.0
    vrms = np.asarray(vrms, dtype=np.float64)
    image0 = np.zeros((nx, ny, nt))
    fold0 = np.zeros((nx, ny, nt), dtype=np.int32)
    if init_nonzero:
        image0 = rng.standard_normal((nx, ny, nt))
        fold0 = rng.integers(0, 5, (nx, ny, nt)).astype(np.int32)

    q, s, t0s, aps = nco._precompute_time_terms(ot, vrms, dip, min_ap, max_ap)
    image = image0.copy()
    fold = fold0.copy()
    nco._migrate_tile_kernel_optimized(
        amp, sx, sy, rx, ry, mx, my, wtr, dt_in, t_start_in, image, fold,
        ox, oy, q, s, t0s, dip, min_ap, max_ap, taper, spreading, obliquity, aps)

    # the kernel the reference's tests call (ms units, helper functions)
    image_b = image0.copy()
    fold_b = fold0.copy()
    nc._migrate_tile_kernel(
        amp, sx, sy, rx, ry, mx, my, wtr, dt_in, t_start_in, image_b, fold_b,
        ox, oy, ot, vrms, dip, min_ap, max_ap, taper, spreading, obliquity, 1)

    np.savez_compressed(
        OUT / f"migrate_{name}.npz",
        amplitudes=amp, source_x=sx, source_y=sy, receiver_x=rx,
        receiver_y=ry, midpoint_x=mx, midpoint_y=my, trace_weights=wtr,
        dt_in_ms=dt_in, t_start_in_ms=t_start_in, x_axis=ox, y_axis=oy,
        t_axis_ms=ot, vrms=vrms, max_dip_deg=dip, min_aperture_m=min_ap,
        max_aperture_m=max_ap, taper_fraction=taper,
        apply_spreading=spreading, apply_obliquity=obliquity,
        image_init=image0, fold_init=fold0,
        t0_half_sq=q, inv_v_sq=s, t0_s=t0s, apertures=aps,
        image=image, fold=fold, image_numba_cpu=image_b, fold_numba_cpu=fold_b)
    print(f"migrate_{name}: |image|max={np.abs(image).max():.4g} "
          f"fold max={fold.max()} "
          f"(numba_cpu kernel diff {np.abs(image - image_b).max():.3g})")


def _interp_vectors() -> None:
    """interpolate_sample (pstm/kernels/interpolation.py:276-318) for every
    method code on an f32-representable trace, at random positions inside the
    migration's sample bounds and at the edges of each method's linear
    fallback, integer and half-integer positions."""
    from pstm.kernels.interpolation import interpolate_sample

    rng = np.random.default_rng(31)
    ns = 41
    trace = rng.standard_normal(ns).astype(np.float32).astype(np.float64)
    special = [0.0, 0.5, 0.999, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0, 7.999, 8.0, 8.5,
               20.0, 20.25, 20.5, 20.75, ns - 9.0, ns - 8.5, ns - 8.001, ns - 6.0, ns - 5.5,
               ns - 5.0, ns - 4.5, ns - 4.0, ns - 3.5, ns - 3.0, ns - 2.5, ns - 2.0, ns - 1.5,
               ns - 1.001]
    pos = np.concatenate([rng.uniform(0.0, ns - 1.0, 300), np.array(special)])
    codes = np.arange(8)
    value = np.array([[interpolate_sample(trace, float(t), int(c)) for t in pos] for c in codes])
    np.savez_compressed(OUT / "interp_methods.npz", trace=trace, pos=pos, reference_code=codes,
                        value=value)
    print("interpolation: ", value[:, :2])


def main() -> None:
    _install_numba_standin()
    sys.path.insert(0, str(REF))
    if "--only" in sys.argv and sys.argv[sys.argv.index("--only") + 1] == "interp":
        _interp_vectors()
        return
    _interp_vectors()
    import pstm.kernels.numba_cpu as nc
    import pstm.kernels.numba_cpu_optimized as nco
    from pstm.algorithm import anisotropy_vti as vti
    from pstm.algorithm import curved_ray as cr
    from pstm.kernels.interpolation import _linear_interp

    _case(nco, nc, "weights_on", 11)
    _case(nco, nc, "weights_off", 12, spreading=False, obliquity=False)
    _case(nco, nc, "spreading_only", 13, obliquity=False, weights=True)
    _case(nco, nc, "taper_dense", 14, max_ap=250.0, min_ap=250.0,
          extent=(-150.0, 430.0), taper=0.3)
    _case(nco, nc, "accumulate", 15, init_nonzero=True, t_start_in=20.0,
          t0_out=8.0)
    _case(nco, nc, "t0_edge", 16, snap=True, dt_in=2.0, n_samples=201,
          dt_out=2.0)
    _case(nco, nc, "const_v_small_ap", 17, vrms=np.full(50, 2000.0),
          max_ap=150.0, min_ap=50.0, dip=30.0, nx=6, ny=10)

    # linear interpolation helper (pstm/kernels/interpolation.py:68-92)
    rng = np.random.default_rng(21)
    trace = rng.standard_normal(37)
    pos = np.concatenate([rng.uniform(-2.0, 39.0, 200),
                          np.array([0.0, 35.0, 35.999, 36.0, 36.5, -1e-9])])
    lin = np.array([_linear_interp(trace, p) for p in pos])
    np.savez_compressed(OUT / "linear_interp.npz", trace=trace, pos=pos,
                        value=lin)

    # curved-ray and VTI DSR traveltimes
    t0 = rng.uniform(0.0, 4.0, 400)
    hs = rng.uniform(0.0, 4000.0, 400)
    hr = rng.uniform(0.0, 4000.0, 400)
    v0 = rng.uniform(1400.0, 2500.0, 400)
    k = rng.uniform(0.0, 1.2, 400)
    k[:20] = 0.0
    k[20:30] = 5e-11
    t0[30:40] = 0.0
    t_cr = np.array([cr.curved_ray_dsr_traveltime(a, b, c, d, e)
                     for a, b, c, d, e in zip(t0, hs, hr, v0, k)])
    z = np.array([cr.time_to_depth_curved(a / 2, d, e)
                  for a, d, e in zip(t0, v0, k)])
    v = rng.uniform(1500.0, 4500.0, 400)
    eta = rng.uniform(0.0, 0.3, 400)
    eta[:20] = 0.0
    t_vti = np.array([vti.vti_dsr_traveltime(a, b, c, d, e)
                      for a, b, c, d, e in zip(t0, hs, hr, v, eta)])
    t_vti_leg = np.array([vti.vti_one_way_time(a / 2, b, d, e)
                          for a, b, d, e in zip(t0, hs, v, eta)])
    np.savez_compressed(OUT / "traveltimes.npz", t0_s=t0, h_s=hs, h_r=hr,
                        v0=v0, k=k, t_curved=t_cr, depth_curved=z, v_nmo=v,
                        eta=eta, t_vti=t_vti, t_vti_leg=t_vti_leg)
    print("traveltimes: curved", t_cr[:3], "vti", t_vti[:3])


if __name__ == "__main__":
    main()
