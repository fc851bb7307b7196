"""CPU tests that pin the oracle before it is trusted as the checker.

Pins, per oracle function:
  * tile migration (straight ray, 1-D velocity, weights, taper, fold,
    accumulation): golden vectors written from the reference's own
    _migrate_tile_kernel_optimized (tests/golden/make_golden.py), plus an
    independent NumPy brute force (other loop order) and invariants;
  * linear interpolation: golden values of the reference's _linear_interp
    (pstm/kernels/interpolation.py:68-92) through a zero-offset migration;
  * 3-D velocity: NumPy brute force (different code, same definition) and the
    reduction to the 1-D case for a laterally constant cube;
  * curved ray: golden values of curved_ray_dsr_traveltime, the k -> 0 and
    h = 0 closed forms, and numerical ray shooting in V(z) = v0 + k z;
  * VTI: golden values of vti_dsr_traveltime / vti_one_way_time and the
    eta = 0 reduction to the straight-ray DSR;
  * triangle anti-aliasing (not in the reference): explicit triangle
    convolution at integer positions, L = 1 equals linear interpolation,
    a constant trace is preserved.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import brute
from oracle import oracle as orc
from pm_amd.kernels import _abi
from tests.conftest import golden_inputs

GOLDEN_CASES = ["weights_on", "weights_off", "spreading_only", "taper_dense", "accumulate",
                "t0_edge", "const_v_small_ap"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_oracle_matches_reference_golden(golden_dir, case):
    g = np.load(golden_dir / f"migrate_{case}.npz")
    inp = golden_inputs(g)
    fold0 = g["fold_init"]
    res = orc.migrate(inp, image_init=g["image_init"], threads=4)
    # fold: per-pillar count broadcast over t, added to the caller's array
    dfold = g["fold"] - fold0
    assert np.all(dfold == dfold[:, :, :1])
    np.testing.assert_array_equal(res.fold2d, dfold[:, :, 0])
    assert res.fold2d.sum() > 0
    scale = np.abs(g["image"]).max()
    assert np.abs(res.image - g["image"]).max() <= 1e-12 * scale
    # the precomputed time terms of the reference (numba_cpu_optimized.py:358-401)
    t = g["t_axis_ms"] / 1000.0
    np.testing.assert_allclose(g["t0_half_sq"], (t * 0.5) ** 2, rtol=0, atol=0)
    np.testing.assert_allclose(g["inv_v_sq"], 1.0 / g["vrms"] ** 2, rtol=1e-15)


def test_reference_kernels_disagree_only_at_t0_edge(golden_dir):
    """numba_cpu.py zeroes the spreading/obliquity weights for t < 1 ms
    (:122-123, :145-146) where the registered optimized kernel skips them
    (numba_cpu_optimized.py:218,225); the oracle follows the registered one."""
    for case in GOLDEN_CASES:
        g = np.load(golden_dir / f"migrate_{case}.npz")
        d = np.abs(g["image"] - g["image_numba_cpu"]).max() / np.abs(g["image"]).max()
        if case == "t0_edge":
            assert d > 1e-3
        else:
            assert d < 1e-12
        np.testing.assert_array_equal(g["fold"], g["fold_numba_cpu"])


def test_linear_interpolation_golden(golden_dir):
    """Zero-offset trace under the pillar: t = t0 exactly, so image[k] is the
    reference's _linear_interp(trace, u_k) with all weights 1."""
    g = np.load(golden_dir / "linear_interp.npz")
    trace, pos, val = g["trace"], g["pos"], g["value"]
    keep = pos >= 0
    pos, val = pos[keep], val[keep]
    dt = 4.0
    amp = trace.astype(np.float32)[None, :]
    z = np.zeros(1)
    inp = orc.MigrateInputs(
        amplitudes=amp, source_x=z, source_y=z, receiver_x=z, receiver_y=z, midpoint_x=z,
        midpoint_y=z, x_axis=z, y_axis=z, t_axis_ms=pos * dt, vrms=np.full(len(pos), 2000.0),
        dt_in_ms=dt, max_aperture_m=100.0, min_aperture_m=100.0, apply_spreading=False,
        apply_obliquity=False)
    res = orc.migrate(inp, threads=1)
    # golden trace is f64, the kernel reads f32 amplitudes: 1e-6 absorbs the cast
    np.testing.assert_allclose(res.image[0, 0], val, rtol=0, atol=1e-6 * np.abs(trace).max())
    assert np.all(res.image[0, 0][pos >= len(trace) - 1] == 0.0)


def _random_inputs(seed, nx=5, ny=6, nt=40, ntr=50, ns=120, vel3d=False, **kw):
    rng = np.random.default_rng(seed)
    mx = rng.uniform(-100, 350, ntr)
    my = rng.uniform(-100, 350, ntr)
    h = rng.uniform(0, 500, ntr)
    az = rng.uniform(0, 2 * np.pi, ntr)
    sx, sy = mx - h / 2 * np.cos(az), my - h / 2 * np.sin(az)
    rx, ry = mx + h / 2 * np.cos(az), my + h / 2 * np.sin(az)
    t = 6.0 * np.arange(nt)
    v = 1700 + 300 * t / 1000
    if vel3d:
        X = np.linspace(0, 1, nx)[:, None, None]
        Y = np.linspace(0, 1, ny)[None, :, None]
        v = v[None, None, :] * (1 + 0.15 * X - 0.1 * Y)
    args = dict(amplitudes=rng.standard_normal((ntr, ns)).astype(np.float32), source_x=sx,
                source_y=sy, receiver_x=rx, receiver_y=ry, midpoint_x=(sx + rx) / 2,
                midpoint_y=(sy + ry) / 2, x_axis=50.0 * np.arange(nx), y_axis=45.0 * np.arange(ny),
                t_axis_ms=t, vrms=v, dt_in_ms=4.0, trace_weights=rng.uniform(0.5, 1.5, ntr),
                max_aperture_m=300.0, min_aperture_m=80.0, max_dip_deg=40.0,
                taper_fraction=0.2)
    args.update(kw)
    return orc.MigrateInputs(**args)


def _brute(inp):
    return brute.migrate_plain(
        inp.amplitudes, inp.source_x, inp.source_y, inp.receiver_x, inp.receiver_y,
        inp.midpoint_x, inp.midpoint_y, inp.trace_weights, inp.x_axis, inp.y_axis,
        inp.t_axis_ms, inp.vrms, inp.dt_in_ms, inp.t_start_in_ms, inp.max_aperture_m,
        inp.min_aperture_m, inp.max_dip_deg, inp.taper_fraction, inp.apply_spreading,
        inp.apply_obliquity)


@pytest.mark.parametrize("vel3d", [False, True])
@pytest.mark.parametrize("seed", [1, 2])
def test_oracle_matches_bruteforce(seed, vel3d):
    inp = _random_inputs(seed, vel3d=vel3d)
    res = orc.migrate(inp, threads=2)
    img, fold = _brute(inp)
    np.testing.assert_array_equal(res.fold2d, fold)
    assert rel(res.image, img) < 1e-12
    assert np.abs(res.image).max() > 0


def test_laterally_constant_cube_equals_1d():
    inp1 = _random_inputs(3)
    inp3 = _random_inputs(3)
    inp3.vrms = np.broadcast_to(inp1.vrms, (5, 6, len(inp1.vrms))).copy()
    a, b = orc.migrate(inp1, threads=2), orc.migrate(inp3, threads=2)
    np.testing.assert_array_equal(a.image, b.image)
    assert a.n_effective == b.n_effective


def test_linearity_zero_and_permutation():
    inp = _random_inputs(4)
    base = orc.migrate(inp, threads=2)
    inp2 = _random_inputs(4)
    inp2.amplitudes = (inp.amplitudes * np.float32(2.0))
    assert rel(orc.migrate(inp2, threads=2).image, 2 * base.image) < 1e-14
    inp0 = _random_inputs(4)
    inp0.amplitudes = np.zeros_like(inp.amplitudes)
    assert np.all(orc.migrate(inp0, threads=2).image == 0)
    perm = np.random.default_rng(0).permutation(inp.amplitudes.shape[0]).astype(np.int32)
    inpp = _random_inputs(4)
    inpp.trace_index = perm
    resp = orc.migrate(inpp, threads=2)
    assert rel(resp.image, base.image) < 1e-13
    np.testing.assert_array_equal(resp.fold2d, base.fold2d)


def test_far_traces_give_zero_fold():
    """tests/test_kernels.py:132-150 of the reference."""
    inp = _random_inputs(5)
    for name in ("source_x", "receiver_x", "midpoint_x"):
        setattr(inp, name, getattr(inp, name) + 10000.0)
    res = orc.migrate(inp, threads=1)
    assert res.fold2d.sum() == 0 and np.all(res.image == 0) and res.n_effective == 0


def test_impulse_response_support():
    """A single spike at sample n0 only reaches output samples whose DSR time
    falls within one input sample of it."""
    inp = _random_inputs(6, ntr=1)
    amp = np.zeros_like(inp.amplitudes)
    n0 = 60
    amp[0, n0] = 1.0
    inp.amplitudes = amp
    res = orc.migrate(inp, threads=1)
    nz = np.argwhere(res.image != 0)
    assert len(nz) > 0
    for ix, iy, k in nz:
        ox, oy = inp.x_axis[ix], inp.y_axis[iy]
        v = inp.vrms[k]
        t0 = inp.t_axis_ms[k] / 1000
        t = (np.sqrt((t0 / 2) ** 2 + ((ox - inp.source_x[0]) ** 2 + (oy - inp.source_y[0]) ** 2) / v ** 2)
             + np.sqrt((t0 / 2) ** 2 + ((ox - inp.receiver_x[0]) ** 2 + (oy - inp.receiver_y[0]) ** 2) / v ** 2))
        assert abs(t / (inp.dt_in_ms / 1000) - n0) < 1.0


# --- traveltimes ----------------------------------------------------------

def test_traveltimes_golden(golden_dir):
    g = np.load(golden_dir / "traveltimes.npz")
    cr = np.array([orc.curved_dsr(*a) for a in zip(g["t0_s"], g["h_s"], g["h_r"], g["v0"], g["k"])])
    np.testing.assert_allclose(cr, g["t_curved"], rtol=1e-13, atol=1e-15)
    vt = np.array([orc.vti_dsr(*a) for a in zip(g["t0_s"], g["h_s"], g["h_r"], g["v_nmo"], g["eta"])])
    np.testing.assert_allclose(vt, g["t_vti"], rtol=1e-13, atol=1e-15)
    leg = np.array([orc.vti_leg(a / 2, b, c, d) for a, b, c, d in
                    zip(g["t0_s"], g["h_s"], g["v_nmo"], g["eta"])])
    np.testing.assert_allclose(leg, g["t_vti_leg"], rtol=1e-13, atol=1e-15)


def test_closed_forms():
    rng = np.random.default_rng(9)
    for _ in range(50):
        t0, hs, hr = rng.uniform(0.05, 3), rng.uniform(0, 3000), rng.uniform(0, 3000)
        v = rng.uniform(1500, 4000)
        dsr = np.sqrt((t0 / 2) ** 2 + hs ** 2 / v ** 2) + np.sqrt((t0 / 2) ** 2 + hr ** 2 / v ** 2)
        assert abs(orc.vti_dsr(t0, hs, hr, v, 0.0) - dsr) < 1e-12 * dsr
        assert abs(orc.curved_dsr(t0, hs, hr, v, 0.0) - dsr) < 1e-12 * dsr
        k = rng.uniform(0.1, 1.0)
        assert abs(orc.curved_dsr(t0, 0.0, 0.0, v, k) - t0) < 1e-12 * t0
        # VTI with eta > 0 is faster than hyperbolic at offset
        assert orc.vti_dsr(t0, hs, hr, v, 0.1) <= dsr + 1e-15


def _shoot_time(h, z, v0, k):
    """Numerical ray in V(z) = v0 + k z: find the ray parameter p that reaches
    horizontal distance h at depth z and integrate its traveltime."""
    from scipy.integrate import quad
    from scipy.optimize import brentq

    def x_of(p):
        return quad(lambda zz: p * (v0 + k * zz) / np.sqrt(1 - (p * (v0 + k * zz)) ** 2), 0, z,
                    epsabs=1e-12, epsrel=1e-12)[0]

    pmax = (1 - 1e-12) / (v0 + k * z)
    p = brentq(lambda p: x_of(p) - h, 0.0, pmax, xtol=1e-16, rtol=1e-14)
    return quad(lambda zz: 1 / ((v0 + k * zz) * np.sqrt(1 - (p * (v0 + k * zz)) ** 2)), 0, z,
                epsabs=1e-13, epsrel=1e-13)[0]


def test_curved_ray_against_ray_shooting():
    """Ray shooting pins the exact circular-ray time arccosh(1 + k^2 (x^2 +
    z^2) / (2 v0 v_z)) / k; the reference's one-way formula
    (pstm/algorithm/curved_ray.py:43-52), which the oracle and the HIP kernel
    follow for drop-in parity (golden vectors above), agrees with it only at
    x = 0 and runs early at offset (about 14 % at x = 1.2 km, z = 2.2 km)."""
    for (t0, h, v0, k) in [(1.0, 300.0, 1600.0, 0.5), (2.0, 1200.0, 1800.0, 0.8),
                          (0.6, 150.0, 2000.0, 0.3)]:
        z = (v0 / k) * (np.exp(k * t0 / 2) - 1)
        vz = v0 + k * z
        with np.errstate(all="ignore"):
            import warnings

            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                t_num = _shoot_time(h, z, v0, k)
        t_exact = np.arccosh(1 + k * k * (h * h + z * z) / (2 * v0 * vz)) / k
        assert abs(t_exact - t_num) < 1e-9
        t_ref = orc.curved_dsr(t0, h, 0.0, v0, k) - t0 / 2  # receiver leg at h=0 is tau
        assert t_ref < t_exact
        assert abs(orc.curved_dsr(t0, 0.0, 0.0, v0, k) - t0) < 1e-12


# --- triangle anti-aliasing -------------------------------------------------

def _p_ext(P, s_end, ns, n):
    if n <= 0:
        return 0.0
    if n <= ns:
        return P[n]
    return P[ns] + (n - ns) * s_end


def test_integrated_second_difference_is_triangle_filter():
    rng = np.random.default_rng(12)
    tr = rng.standard_normal(40).astype(np.float32)
    P, S = orc.integrate_traces(tr[None, :])
    P, s_end = P[0], S[0]
    for L in (1, 2, 3, 5):
        for n in range(-3, 44):
            d2 = (_p_ext(P, s_end, 40, n + L) - 2 * _p_ext(P, s_end, 40, n)
                  + _p_ext(P, s_end, 40, n - L)) / L ** 2
            assert abs(d2 - brute.triangle_filter(tr, n, L)) < 1e-9, (L, n)


def test_triangle_aa_reduces_to_linear_when_dip_is_zero():
    """grid spacing 0 -> L clamps to 1 -> plain linear interpolation."""
    inp = _random_inputs(7)
    inp_aa = _random_inputs(7, aa_mode=_abi.AA_TRIANGLE, grid_dx=0.0, grid_dy=0.0)
    a, b = orc.migrate(inp, threads=2), orc.migrate(inp_aa, threads=2)
    assert rel(b.image, a.image) < 1e-9
    assert a.n_effective == b.n_effective


def test_triangle_aa_smooths_and_preserves_constant():
    inp = _random_inputs(8, aa_mode=_abi.AA_TRIANGLE, grid_dx=40.0, grid_dy=40.0)
    inp.amplitudes = np.full_like(inp.amplitudes, 0.5)
    inp_plain = _random_inputs(8)
    inp_plain.amplitudes = inp.amplitudes
    a, b = orc.migrate(inp_plain, threads=2), orc.migrate(inp, threads=2)
    # a constant trace is unchanged by the triangle except within L of the ends
    inner = np.abs(a.image) > 0
    r = np.abs(b.image[inner] - a.image[inner]) / np.abs(a.image[inner])
    assert np.median(r) < 1e-9
    # random traces: AA attenuates (the filter is a low-pass)
    inp_r = _random_inputs(8, aa_mode=_abi.AA_TRIANGLE, grid_dx=40.0, grid_dy=40.0)
    assert np.linalg.norm(orc.migrate(inp_r, threads=2).image) < np.linalg.norm(
        orc.migrate(_random_inputs(8), threads=2).image)


def test_point_diffractor_focuses():
    """Known answer (the analogue of tests/test_integration.py:121-130): the
    point-diffractor job focuses at its bin and t0 = 800 ms."""
    from pm_amd.synthetic import point_diffractor_job

    job = point_diffractor_job(n_traces=1024)
    sel = slice(24, 40)
    inp = orc.MigrateInputs(
        amplitudes=job.amplitudes.numpy(), source_x=job.source_x, source_y=job.source_y,
        receiver_x=job.receiver_x, receiver_y=job.receiver_y, midpoint_x=job.midpoint_x,
        midpoint_y=job.midpoint_y, x_axis=job.x_axis[sel], y_axis=job.y_axis[sel],
        t_axis_ms=job.t_axis_ms, vrms=job.vrms_1d, dt_in_ms=job.dt_ms,
        max_aperture_m=1000.0, min_aperture_m=1000.0)
    res = orc.migrate(inp)
    ix, iy, k = np.unravel_index(np.argmax(res.image), res.image.shape)
    assert (job.x_axis[sel][ix], job.y_axis[sel][iy]) == (1000.0, 1000.0)
    assert abs(job.t_axis_ms[k] - 800.0) <= 2.0
