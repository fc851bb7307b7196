"""CPU pins of the oracle paths that the reference's golden vectors leave open.

* Curved-ray migration (oracle loop over per-sample depth z(t0 / 2) and the
  spreading velocity v(z)): k = 0 equals straight-ray migration in v0, image
  and fold; the impulse response of a single spike lies where the
  golden-pinned scalar time curved_dsr (curved_ray.py:287-316) is within one
  sample of the spike; the spreading velocity of a constant trace is
  v0 exp(k t0 / 2), the velocity at the depth whose vertical one-way time is
  t0 / 2 in v(z) = v0 + k z (the integral of dz / v(z)).
* Triangle anti-aliasing with fractional half width L: continuous in L across
  every integer from both sides, second moment of the effective filter equal
  to the closed form (L^2 - 1) / 6 at integer L and increasing with L in
  between (a blend with its weights swapped fails both), and the migration
  with anti-aliasing equal to an independent brute force (finite-difference
  operator dips, explicit triangle convolutions).
* 3-D velocity: VelocityCube.resample equals SciPy's
  RegularGridInterpolator(method="linear") on coordinates clamped to the
  cube, as pstm/data/velocity_model.py:375-396 builds it.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import brute
from oracle import oracle as orc
from pm_amd.kernels import _abi


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _inputs(seed, ntr=40, ns=160, nx=5, ny=4, nt=60, v=None, **kw):
    rng = np.random.default_rng(seed)
    mx = rng.uniform(-100, 300, ntr)
    my = rng.uniform(-100, 250, ntr)
    h = rng.uniform(0, 600, ntr)
    az = rng.uniform(0, 2 * np.pi, ntr)
    sx, sy = mx - h / 2 * np.cos(az), my - h / 2 * np.sin(az)
    rx, ry = mx + h / 2 * np.cos(az), my + h / 2 * np.sin(az)
    t = 6.0 * np.arange(nt)
    if v is None:
        v = 1700 + 300 * t / 1000
    args = dict(amplitudes=rng.standard_normal((ntr, ns)).astype(np.float32), source_x=sx,
                source_y=sy, receiver_x=rx, receiver_y=ry, midpoint_x=(sx + rx) / 2,
                midpoint_y=(sy + ry) / 2, x_axis=50.0 * np.arange(nx), y_axis=45.0 * np.arange(ny),
                t_axis_ms=t, vrms=np.broadcast_to(v, (nt,)).astype(float).copy(), dt_in_ms=4.0,
                trace_weights=rng.uniform(0.5, 1.5, ntr), max_aperture_m=400.0, min_aperture_m=120.0,
                max_dip_deg=40.0, taper_fraction=0.2)
    args.update(kw)
    return orc.MigrateInputs(**args)


# --- curved ray ---------------------------------------------------------------

def test_curved_k0_migration_equals_straight_ray_in_v0():
    """k = 0 (below the reference's 1e-10 switch, curved_ray.py:35-38):
    z = v0 t0 / 2, t = sqrt(h^2 + z^2) / v0 per leg and the spreading
    velocity v0 — the straight-ray migration in v_rms = v0.  A depth taken
    from t0 instead of t0 / 2 or a leg time from z = v0 t0 fails this."""
    v0 = 1850.0
    st = _inputs(3, v=v0)
    cr = _inputs(3, v=v0, traveltime=_abi.TT_CURVED, curved_v0=v0, curved_k=0.0)
    a, b = orc.migrate(st, threads=2), orc.migrate(cr, threads=2)
    np.testing.assert_array_equal(a.fold2d, b.fold2d)
    assert np.abs(a.image).max() > 0
    assert rel(b.image, a.image) < 1e-12
    assert a.n_effective == b.n_effective


@pytest.mark.parametrize("k", [0.4, 0.9])
def test_curved_impulse_response_follows_the_pinned_time(k):
    """A single spike at sample n0 reaches exactly the output samples whose
    curved-ray time (curved_dsr, golden-pinned against the reference) lies
    within one input sample of n0, and every such sample inside the aperture
    and the trace receives a non-zero value."""
    v0 = 1600.0
    inp = _inputs(5, ntr=1, ns=400, nx=6, ny=6, nt=120, v=3000.0, traveltime=_abi.TT_CURVED,
                  curved_v0=v0, curved_k=k, max_aperture_m=1500.0, min_aperture_m=1500.0)
    n0 = 150
    amp = np.zeros_like(inp.amplitudes)
    amp[0, n0] = 1.0
    inp.amplitudes = amp
    res = orc.migrate(inp, threads=1)
    dt = inp.dt_in_ms / 1000.0
    hits = 0
    for ix, ox in enumerate(inp.x_axis):
        for iy, oy in enumerate(inp.y_axis):
            hs = np.hypot(ox - inp.source_x[0], oy - inp.source_y[0])
            hr = np.hypot(ox - inp.receiver_x[0], oy - inp.receiver_y[0])
            for kk, tms in enumerate(inp.t_axis_ms):
                u = orc.curved_dsr(tms / 1000.0, hs, hr, v0, k) / dt
                val = res.image[ix, iy, kk]
                if abs(u - n0) >= 1.0:
                    assert val == 0.0, (ix, iy, kk, u)
                elif abs(u - n0) < 0.999 and tms > 2.0:
                    assert val != 0.0, (ix, iy, kk, u)
                    hits += 1
    assert hits > 10


def test_curved_spreading_uses_velocity_at_depth():
    """Constant trace, one trace, weights on: image(k) = 1 / (v_w t) * t0 / t,
    so image t^2 / t0 = 1 / v_w.  The reference's spreading velocity is the
    velocity at the image depth (curved_ray.py:268-270); in v(z) = v0 + k z
    the depth whose vertical one-way time is t0 / 2 has v = v0 exp(k t0 / 2).
    A spreading velocity v0 fails this."""
    v0, k = 1700.0, 0.6
    inp = _inputs(7, ntr=1, ns=800, nx=3, ny=3, nt=150, v=3000.0, traveltime=_abi.TT_CURVED,
                  curved_v0=v0, curved_k=k, max_aperture_m=3000.0, min_aperture_m=3000.0,
                  taper_fraction=1e-9, trace_weights=np.ones(1))
    inp.source_x[:] = inp.receiver_x[:] = 40.0
    inp.source_y[:] = inp.receiver_y[:] = 30.0
    inp.midpoint_x[:] = 40.0
    inp.midpoint_y[:] = 30.0
    inp.amplitudes = np.ones_like(inp.amplitudes)
    res = orc.migrate(inp, threads=1)
    checked = 0
    for ix, ox in enumerate(inp.x_axis):
        for iy, oy in enumerate(inp.y_axis):
            h = np.hypot(ox - 40.0, oy - 30.0)
            for kk in range(1, 150):
                t0 = inp.t_axis_ms[kk] / 1000.0
                t = orc.curved_dsr(t0, h, h, v0, k)
                if t <= 0.001 or t / 0.004 >= 799:
                    continue
                got = res.image[ix, iy, kk] * t * t / t0
                assert got == pytest.approx(1.0 / (v0 * np.exp(k * t0 / 2)), rel=1e-10)
                checked += 1
    assert checked > 500


# --- triangle anti-aliasing ----------------------------------------------------

def _trace(seed=11, n=80):
    return np.random.default_rng(seed).standard_normal(n).astype(np.float32)


def test_triangle_is_continuous_in_half_width():
    """At every integer L the filter of L - e and of L + e agree with the
    integer-L triangle (to O(e)); a blend with its weights swapped jumps from
    the L - 1 to the L + 1 triangle there."""
    tr = _trace()
    eps = 1e-7
    for n in (2, 3, 4, 7):
        for u in (30.0, 31.25, 40.6):
            mid = orc.triangle_at(tr, u, float(n))
            lo = orc.triangle_at(tr, u, n - eps)
            hi = orc.triangle_at(tr, u, n + eps)
            assert abs(lo - mid) < 1e-5 and abs(hi - mid) < 1e-5, (n, u, lo, mid, hi)
            # and the integer-L value is the explicit triangle (lerp in u)
            i0 = int(np.floor(u))
            f = u - i0
            want = (1 - f) * brute.triangle_filter(tr, i0, n) + f * brute.triangle_filter(tr, i0 + 1, n)
            assert abs(mid - want) < 1e-9


def _moments(L, n0=40, n=81):
    """Effective filter taps of triangle_at at integer position n0: response
    to a unit spike at n0 + m."""
    taps = np.zeros(n)
    for j in range(n):
        tr = np.zeros(n, np.float32)
        tr[j] = 1.0
        taps[j] = orc.triangle_at(tr, float(n0), L)
    m = np.arange(n) - n0
    return taps.sum(), (taps * m).sum(), (taps * m * m).sum()


def test_triangle_second_moment_grows_with_half_width():
    """Unit DC gain, zero first moment, second moment (L^2 - 1) / 6 at integer
    L (the discrete triangle sum_{|m|<L} (L - |m|) m^2 / L^2) and strictly
    increasing with L between the integers."""
    prev = -1.0
    for L in np.arange(1.0, 6.01, 0.125):
        dc, m1, m2 = _moments(float(L))
        assert dc == pytest.approx(1.0, abs=1e-12)
        assert abs(m1) < 1e-12
        if float(L).is_integer():
            assert m2 == pytest.approx((L * L - 1) / 6, abs=1e-12)
        assert m2 > prev, L
        prev = m2


@pytest.mark.parametrize("seed", [21, 22])
def test_triangle_aa_migration_matches_bruteforce(seed):
    """Migration with triangle anti-aliasing against the NumPy brute force
    that takes the operator dip from finite differences of the traveltime and
    applies the triangles as explicit convolutions (no integrated traces)."""
    kw = dict(aa_mode=_abi.AA_TRIANGLE, grid_dx=30.0, grid_dy=25.0, aa_max_half_width=6)
    inp = _inputs(seed, ntr=24, ns=200, nx=4, ny=3, nt=70, **kw)
    res = orc.migrate(inp, threads=2)
    img, fold = brute.migrate_plain(
        inp.amplitudes, inp.source_x, inp.source_y, inp.receiver_x, inp.receiver_y,
        inp.midpoint_x, inp.midpoint_y, inp.trace_weights, inp.x_axis, inp.y_axis,
        inp.t_axis_ms, inp.vrms, inp.dt_in_ms, inp.t_start_in_ms, inp.max_aperture_m,
        inp.min_aperture_m, inp.max_dip_deg, inp.taper_fraction, aa=(30.0, 25.0, 6))
    np.testing.assert_array_equal(res.fold2d, fold)
    assert np.abs(img).max() > 0
    assert rel(res.image, img) < 1e-7
    plain = orc.migrate(_inputs(seed, ntr=24, ns=200, nx=4, ny=3, nt=70), threads=2)
    assert rel(plain.image, img) > 1e-2   # the filter does act on this case


# --- 3-D velocity resampling ----------------------------------------------------

def test_velocity_cube_resample_matches_scipy_regular_grid_interpolator():
    from scipy.interpolate import RegularGridInterpolator

    from pm_amd.synthetic import lateral_cube, vti_model

    rng = np.random.default_rng(3)
    for cube in (lateral_cube(), vti_model()[0], vti_model()[1]):
        x = np.sort(rng.uniform(cube.x_axis[0] - 300, cube.x_axis[-1] + 300, 23))
        y = np.sort(rng.uniform(cube.y_axis[0] - 300, cube.y_axis[-1] + 300, 19))
        t = np.sort(rng.uniform(-50, cube.t_axis_ms[-1] + 400, 31))
        got = cube.resample(x, y, t).numpy()
        rgi = RegularGridInterpolator((cube.x_axis, cube.y_axis, cube.t_axis_ms), cube.data,
                                      method="linear", bounds_error=False, fill_value=None)
        X, Y, T = np.meshgrid(np.clip(x, cube.x_axis[0], cube.x_axis[-1]),
                              np.clip(y, cube.y_axis[0], cube.y_axis[-1]),
                              np.clip(t, cube.t_axis_ms[0], cube.t_axis_ms[-1]), indexing="ij")
        want = rgi(np.stack([X.ravel(), Y.ravel(), T.ravel()], 1)).reshape(got.shape)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-9)
        # grid nodes reproduce the cube
        node = cube.resample(cube.x_axis[::7], cube.y_axis[::5], cube.t_axis_ms[::11]).numpy()
        np.testing.assert_allclose(node, cube.data[::7, ::5, ::11], rtol=1e-13)
