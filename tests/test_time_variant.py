"""Time-variant output sampling: the window plan (CPU, against hand-derived
cases and the oracle's restatement), the HIP resampling kernel against
np.interp, and time-variant migration against the uniform migration and the
FP64 oracle on the window axis (GPU)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import time_variant as otv
from pm_amd import time_variant as tv


# --------------------------------------------------------------------------- CPU


def test_nearest_power_of_2_and_factor_by_hand():
    # ties go to the lower power: 3 -> 2, 6 -> 4, 12 -> 8
    assert [tv.nearest_power_of_2(n) for n in (0, 1, 2, 3, 5, 6, 7, 12, 13)] == [1, 1, 2, 2, 4, 4, 8, 8, 16]
    # Nyquist 1000 / (2 f) over base dt, truncated: 80 Hz at 2 ms -> 3 -> 2
    assert tv.downsample_factor(80.0, 2.0) == 2
    assert tv.downsample_factor(250.0, 2.0) == 1          # 1 -> 1
    assert tv.downsample_factor(10.0, 2.0) == 8           # 25 -> 32, clamped to 8
    assert tv.downsample_factor(10.0, 2.0, max_factor=16) == 16    # 25 -> 32, clamped to 16
    assert tv.downsample_factor(1000.0, 2.0, min_factor=2) == 2   # 0 -> 1, clamped up


def test_constant_table_windows_by_hand():
    """25 Hz everywhere at 2 ms: factor 8, dt 16 ms.  Over [0, 100) ms the
    first window holds 6 samples (0 .. 80) and ends at 96; the remainder
    [96, 100) gets one more sample at 96."""
    table = tv.FrequencyTimeTable([0.0, 5000.0], [25.0, 25.0])
    w = tv.compute_time_windows(0.0, 100.0, 2.0, table)
    assert [(x.t_start_ms, x.t_end_ms, x.dt_effective_ms, x.downsample_factor, x.sample_start, x.sample_end)
            for x in w] == [(0.0, 96.0, 16.0, 8, 0, 6), (96.0, 100.0, 16.0, 8, 6, 7)]
    np.testing.assert_array_equal(tv.sample_times(w), [0, 16, 32, 48, 64, 80, 96])


def test_table_interpolation_end_values_and_nodes():
    table = tv.FrequencyTimeTable.default()
    assert table.fmax(-5.0) == 80.0 and table.fmax(9000.0) == 15.0
    for t, f in zip(table.times_ms, table.frequencies_hz):
        assert table.fmax(t) == pytest.approx(f, abs=1e-12)
    # monotone table: interpolated values stay between the neighbours
    for t in np.linspace(1, 4999, 97):
        i = np.searchsorted(table.times_ms, t) - 1
        lo, hi = sorted(table.frequencies_hz[i:i + 2])
        assert lo - 1e-12 <= table.fmax(t) <= hi + 1e-12
    with pytest.raises(ValueError):
        tv.FrequencyTimeTable([0.0, 0.0], [10.0, 20.0])
    with pytest.raises(ValueError):
        tv.FrequencyTimeTable([0.0, 1.0], [10.0, -1.0])


def test_windows_tile_the_axis():
    table = tv.FrequencyTimeTable.from_pairs(tv.CONFIG_DEFAULT_TABLE)
    w = tv.compute_time_windows(0.0, 1998.0, 2.0, table)
    assert w[0].t_start_ms == 0.0 and w[-1].t_end_ms == 1998.0
    for a, b in zip(w, w[1:]):
        assert a.t_end_ms == b.t_start_ms and a.sample_end == b.sample_start
    assert all(x.downsample_factor in (1, 2, 4, 8) and x.dt_effective_ms == 2.0 * x.downsample_factor for x in w)
    t = tv.sample_times(w)
    assert np.all(np.diff(t) > 0) and t[-1] < 1998.0
    # every window sample lies on the 2 ms axis
    k = tv.uniform_index(t, 2.0 * np.arange(1000))
    np.testing.assert_array_equal(2.0 * k, t)
    assert 3.0 < tv.estimate_speedup(0.0, 1998.0, 2.0, table) < 3.5


@pytest.mark.parametrize("seed", range(6))
def test_windows_match_oracle_on_random_tables(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 7))
    times = np.sort(rng.choice(np.arange(0, 6000, 50), n, replace=False)).astype(float)
    freqs = rng.uniform(8.0, 120.0, n)
    base = float(rng.choice([1.0, 2.0, 4.0]))
    t_max = float(rng.integers(200, 4000))
    lo, hi = 1, int(rng.choice([4, 8, 16]))
    got = tv.compute_time_windows(0.0, t_max, base, tv.FrequencyTimeTable(list(times), list(freqs)), lo, hi)
    want = otv.windows(0.0, t_max, base, list(times), list(freqs), lo, hi)
    assert [(w.t_start_ms, w.t_end_ms, w.dt_effective_ms, w.downsample_factor, w.sample_start, w.sample_end,
             w.f_max_hz) for w in got] == want
    np.testing.assert_array_equal(tv.sample_times(got), otv.tv_times(want))


def test_oracle_resample_linear_data_and_hold():
    """A linear trace is reproduced exactly between window samples; beyond the
    last window sample the last value holds; one sample leaves zeros."""
    w = [(0.0, 40.0, 8.0, 4, 0, 5)]
    xp = otv.tv_times(w)                       # 0, 8, .., 32
    img = (3.0 + 0.5 * xp)[None, None, :]
    out = otv.resample(img, w, 2.0, 20)        # 0 .. 38 ms
    np.testing.assert_allclose(out[0, 0, :17], 3.0 + 0.5 * 2.0 * np.arange(17), rtol=0, atol=1e-12)
    np.testing.assert_array_equal(out[0, 0, 17:], 3.0 + 0.5 * 32.0)
    np.testing.assert_array_equal(otv.resample(img[..., :1], [(0.0, 8.0, 8.0, 4, 0, 1)], 2.0, 5), 0.0)


def test_uniform_index_rejects_off_axis_samples():
    with pytest.raises(ValueError):
        tv.uniform_index(np.array([0.0, 3.0]), 2.0 * np.arange(10))
    with pytest.raises(ValueError):
        tv.uniform_index(np.array([0.0, 40.0]), 2.0 * np.arange(10))


# --------------------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def kernel():
    from pm_amd.kernels import HipKernel, KernelConfig

    k = HipKernel()
    k.initialize(KernelConfig())
    yield k
    k.cleanup()


@pytest.mark.gpu
@pytest.mark.parametrize("n_tv", [1, 2, 37, 308, 5000])
def test_resample_kernel_bit_exact_with_np_interp(kernel, n_tv):
    """(5000 nodes: more than the kernel copies to LDS, so it searches the
    axis in global memory)"""
    import torch

    rng = np.random.default_rng(n_tv)
    xp = np.sort(rng.choice(np.arange(0, max(2000, 4 * n_tv), 2.0), n_tv, replace=False)) + 4.0
    img = rng.normal(size=(7, 5, n_tv)).astype(np.float32)
    x = np.concatenate([[-3.0, xp[0]], np.arange(1000) * 2.0, [xp[-1], xp[-1] + 1.0]])
    got = kernel.resample_uniform(torch.as_tensor(img, device="cuda"), xp, x).cpu().numpy()
    if n_tv < 2:
        np.testing.assert_array_equal(got, 0.0)
        return
    want = np.stack([np.interp(x, xp, row.astype(np.float64)) for row in img.reshape(-1, n_tv)]).astype(np.float32)
    np.testing.assert_array_equal(got.reshape(-1, x.size), want)


@pytest.mark.gpu
def test_resample_rejects_bad_axes(kernel):
    import torch

    img = torch.zeros((2, 3, 4), device="cuda")
    with pytest.raises(ValueError):
        kernel.resample_uniform(img, np.array([0.0, 1.0, 2.0]), np.arange(5.0))
    with pytest.raises(ValueError):
        kernel.resample_uniform(img, np.array([0.0, 2.0, 1.0, 3.0]), np.arange(5.0))


def _job():
    from pm_amd.synthetic import reflector_job

    return reflector_job(seed=5, n_traces=6000, n_samples=601, n_out=24, nt=500, d_out=25.0, tile=12,
                         n_planes=6, n_diffractors=10, antialias="none", device="cuda")


@pytest.mark.gpu
def test_time_variant_equals_uniform_at_window_samples_and_oracle(kernel, tmp_path):
    """Time-variant migration equals the uniform migration at the window
    samples (same gather, same per-sample terms), its fold is the uniform
    fold, one tile is within the FP32 bar of the FP64 oracle on the window
    axis, and the uniform resampling is np.interp of the window image."""
    import torch

    from oracle import oracle as orc
    from pm_amd.pipeline.executor import TileExecutor

    job = _job()
    table = tv.FrequencyTimeTable.from_pairs(tv.CONFIG_DEFAULT_TABLE)
    w = tv.compute_time_windows(float(job.t_axis_ms[0]), float(job.t_axis_ms[-1]), 2.0, table)
    k = tv.uniform_index(tv.sample_times(w), job.t_axis_ms)
    assert len(k) < len(job.t_axis_ms) / 2
    uni = TileExecutor(kernel, job)
    uni.plan().run()
    u_img, u_fold = uni.assemble()
    ex = TileExecutor(kernel, job, time_windows=w)
    assert ex.shape == (24, 24, len(k))
    ex.plan().run()
    t_img, t_fold = ex.assemble()
    torch.cuda.synchronize()
    assert torch.equal(t_fold, u_fold)
    ref = u_img[..., torch.as_tensor(k, device="cuda")]
    rel = float((t_img.double() - ref.double()).norm() / ref.double().norm())
    print("time-variant vs uniform at window samples", rel, bool(torch.equal(t_img, ref)))
    assert rel < 1e-6
    # one tile against the oracle on the window axis
    t = max(ex.tiles, key=lambda t: int(t.rows.shape[0]))
    rows = t.rows.cpu().numpy().astype(np.int64)
    h = ex.host_traces(rows)
    cfg = ex.config
    inp = orc.MigrateInputs(
        amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
        receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
        midpoint_y=h["midpoint_y"], x_axis=t.x_axis.cpu().numpy(), y_axis=t.y_axis.cpu().numpy(),
        t_axis_ms=tv.sample_times(w), vrms=t.vrms.cpu().numpy(), dt_in_ms=job.dt_ms,
        max_aperture_m=cfg.max_aperture_m, min_aperture_m=cfg.min_aperture_m, max_dip_deg=cfg.max_dip_degrees,
        taper_fraction=cfg.taper_fraction, apply_spreading=cfg.apply_spreading,
        apply_obliquity=cfg.apply_obliquity, grid_dx=cfg.grid_dx, grid_dy=cfg.grid_dy)
    o = orc.migrate(inp)
    got = t_img[t.x0:t.x1, t.y0:t.y1].double().cpu().numpy()
    rel_o = np.linalg.norm(got - o.image) / np.linalg.norm(o.image)
    print("time-variant tile vs FP64 oracle", rel_o)
    assert rel_o <= 1e-5
    np.testing.assert_array_equal(t_fold[t.x0:t.x1, t.y0:t.y1].cpu().numpy(), o.fold2d)
    # back on the uniform axis: np.interp of the window image, bit for bit
    full = ex.assemble_uniform().cpu().numpy()
    wins = [(x.t_start_ms, x.t_end_ms, x.dt_effective_ms, x.downsample_factor, x.sample_start, x.sample_end,
             x.f_max_hz) for x in w]
    want = otv.resample(t_img.cpu().numpy(), wins, 2.0, len(job.t_axis_ms)).astype(np.float32)
    np.testing.assert_array_equal(full, want)
    # finalize writes the uniform-axis image
    from pm_amd.io.zarr_v2 import read_zarr_v2

    ex.finalize(tmp_path, chunk_x=12, chunk_y=12)
    vol, attrs = read_zarr_v2(tmp_path / "migrated_stack.zarr")
    np.testing.assert_array_equal(vol, full)
    assert attrs["dt_ms"] == 2.0 and vol.shape[2] == len(job.t_axis_ms)


@pytest.mark.gpu
def test_drop_in_time_variant_tile(kernel):
    """migrate_tile_time_variant with host arrays: the window image replaces
    output.image, the fold is added along t, and resample_to_uniform brings
    it back to the uniform axis."""
    from pm_amd.kernels.base import create_output_tile, create_trace_block, VelocitySlice

    job = _job()
    rows = np.arange(2000)
    tb = create_trace_block(job.amplitudes[rows].cpu().numpy(), job.source_x[rows], job.source_y[rows],
                            job.receiver_x[rows], job.receiver_y[rows], sample_rate_ms=job.dt_ms)
    x, y = job.x_axis[:8], job.y_axis[:6]

    def tile():
        return create_output_tile(x[0], x[-1], 25.0, y[0], y[-1], 25.0, 0.0, float(job.t_axis_ms[-1]), 2.0)

    out = tile()
    assert out.image.shape == (8, 6, len(job.t_axis_ms))
    v = VelocitySlice(vrms=np.full(len(job.t_axis_ms), 2200.0), is_1d=True, t_axis_ms=job.t_axis_ms)
    table = tv.FrequencyTimeTable.default()
    w = tv.compute_time_windows(0.0, float(job.t_axis_ms[-1]), 2.0, table)
    m = kernel.migrate_tile_time_variant(tb, out, v, job.config, w)
    n_tv = len(tv.sample_times(w))
    assert out.image.shape == (8, 6, n_tv) and out.image.dtype == np.float64
    assert m.n_samples_output == 8 * 6 * n_tv
    uni = tile()
    kernel.migrate_tile(tb, uni, v, job.config)
    k = tv.uniform_index(tv.sample_times(w), job.t_axis_ms)
    rel = np.linalg.norm(out.image - uni.image[..., k]) / np.linalg.norm(uni.image[..., k])
    assert rel < 1e-6
    np.testing.assert_array_equal(out.fold, uni.fold)
    back = tv.resample_to_uniform(out.image, w, 2.0, len(job.t_axis_ms), kernel)
    assert back.shape == uni.image.shape
    wins = [(x_.t_start_ms, x_.t_end_ms, x_.dt_effective_ms, x_.downsample_factor, x_.sample_start,
             x_.sample_end, x_.f_max_hz) for x_ in w]
    np.testing.assert_array_equal(back, otv.resample(out.image.astype(np.float32), wins, 2.0,
                                                     len(job.t_axis_ms)).astype(np.float32).astype(np.float64))


@pytest.mark.gpu
def test_time_variant_vti_equals_uniform_at_window_samples(kernel):
    """A VTI job with 3-D velocity and eta: the window samples take the eta of
    their uniform samples, so the time-variant image is the uniform image
    there."""
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import vti_diffractor_job

    job = vti_diffractor_job(seed=6, n_traces=4000, n_samples=801, n_out=20, nt=600, d_out=25.0, tile=10,
                             n_diffractors=12, device="cuda")
    w = tv.compute_time_windows(0.0, float(job.t_axis_ms[-1]), 2.0, tv.FrequencyTimeTable.default())
    k = torch.as_tensor(tv.uniform_index(tv.sample_times(w), job.t_axis_ms), device="cuda")
    uni = TileExecutor(kernel, job)
    uni.plan().run()
    u_img, u_fold = uni.assemble()
    ex = TileExecutor(kernel, job, time_windows=w)
    assert ex.tiles[0].eta is not None and ex.tiles[0].eta.shape[-1] == len(k)
    ex.plan().run()
    t_img, t_fold = ex.assemble()
    assert torch.equal(t_fold, u_fold)
    ref = u_img[..., k]
    rel = float((t_img.double() - ref.double()).norm() / ref.double().norm())
    print("VTI time-variant vs uniform at window samples", rel)
    assert rel < 1e-6
    with pytest.raises(ValueError, match="uniform output axis"):
        ex.run(surrogate=object())
