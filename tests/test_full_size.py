"""The benchmark configurations at their full sizes on the GPU.

Each job is built with its full recipe (pm_amd.synthetic defaults) and run
through the HIP path; the fold is compared bit for bit with the oracle's f64
count on EVERY pillar migrated, and the image with the FP64 oracle on a
deterministic pillar sample of the busiest tile (consecutive pillars of its
middle inline: all that tile's traces, all samples) to relative L2 2e-5 at
1501 and 3e-5 at 2001 samples per trace (the f32 sample-position floor, see
F32_FLOOR_*; the 1e-5 bar holds at 1001 samples, tests/test_gpu_parity.py).

* reflectors: 200,000 traces x 1501 samples, 3-D v_rms, triangle
  anti-aliasing, 256 x 256 x 1000 output in 16 tiles, one plan launch.
* bins: 40 common-offset bins with lognormal counts (1.6M traces x 1501),
  no anti-aliasing, every tile and bin in one persistent, graph-replayed
  launch into one shared image.
* VTI: 500,000 traces x 2001 samples, 3-D v_rms and eta, 512 x 512 x 1500
  output (the whole job resident: traces, velocity and eta volumes, image,
  fold and every tile's trace list); migrated: the stated central region of
  32 x 32 pillars, exact VTI against the oracle and the learned traveltime
  (f32-trained network, bf16 weights, the growing t0 lattice of the
  benchmark, pm_amd.lattice.SPARSE) against exact to image relative L2
  1e-3, with its per-tile held-out check.

* streamed (the largest single-GPU job): 60 common-offset bins, ~2.4M
  traces x 2001 samples (19.2 GB) written as Zarr v2 + Parquet, evicted from
  the page cache and streamed to HBM bin by bin while each bin migrates into
  a 1024 x 1024 x 2000 image (8.4 GB) with the learned VTI traveltime on the
  stated central 32 x 32 pillars; fold against the oracle on every pillar of
  the region, image against the exact migration of the streamed traces.

Progress lines name the job and step (each job takes tens of seconds).
"""

from __future__ import annotations

import time

import numpy as np
import pytest

from oracle import oracle as orc
from pm_amd.kernels import _abi

pytestmark = pytest.mark.gpu

# Image bars at full trace length.  The gather computes sample positions
# u = t / dt in f32 (traveltimes on the FP32 VALU): their error is
# about 1e-7 u samples, and for the 30 Hz events of these jobs a position
# error e costs about 2 pi 30 Hz dt e of an amplitude.  At u ~ 1000 (the
# scaled and 1001-sample tests) that is below 1e-5 of the image; at 1501 and
# 2001 samples it reaches 1-2e-5, measured and stated here rather than hidden
# behind shorter traces.
F32_FLOOR_1501 = 2e-5
F32_FLOOR_2001 = 3e-5


def say(msg: str) -> None:
    print(f"[full-size] {msg}", flush=True)


@pytest.fixture(scope="module")
def kernel():
    import torch

    from pm_amd.kernels import HipKernel, KernelConfig

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    k = HipKernel()
    k.initialize(KernelConfig())
    yield k
    k.cleanup()


def _fold_check(ex, tiles):
    """Every pillar of the listed tiles against the oracle's f64 fold."""
    tr = ex.traces
    mx, my = tr.midpoint_x.cpu().numpy(), tr.midpoint_y.cpu().numpy()
    for i in tiles:
        t = ex.tiles[i]
        want = orc.fold(t.x_axis.cpu().numpy(), t.y_axis.cpu().numpy(), mx, my, ex.config.max_aperture_m,
                        trace_index=t.rows.cpu().numpy())
        np.testing.assert_array_equal(t.fold.cpu().numpy(), want)
    # the tile lists hold every trace within the aperture of their pillars:
    # the fold over ALL traces is the same
    t = ex.tiles[tiles[0]]
    allf = orc.fold(t.x_axis.cpu().numpy(), t.y_axis.cpu().numpy(), mx, my, ex.config.max_aperture_m)
    np.testing.assert_array_equal(t.fold.cpu().numpy(), allf)


def _image_sample(ex, tile_index: int, n_pillars: int, traveltime: int, eta: bool = False):
    """The oracle on n_pillars consecutive pillars of the middle inline of one
    tile (all that tile's traces, all samples), against the HIP image of the
    same pillars."""
    t = ex.tiles[tile_index]
    rows = t.rows.cpu().numpy().astype(np.int64)
    h = ex.host_traces(rows)
    cfg = ex.config
    nx, ny = t.x1 - t.x0, t.y1 - t.y0
    sl = (slice(nx // 2, nx // 2 + 1), slice(ny // 2 - n_pillars // 2, ny // 2 - n_pillars // 2 + n_pillars))
    v = t.vrms.cpu().numpy()
    inp = orc.MigrateInputs(
        amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"], receiver_x=h["receiver_x"],
        receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"], midpoint_y=h["midpoint_y"],
        x_axis=t.x_axis.cpu().numpy()[sl[0]], y_axis=t.y_axis.cpu().numpy()[sl[1]],
        t_axis_ms=ex.t_axis.cpu().numpy(), vrms=v[sl] if v.ndim == 3 else v,
        eta=t.eta.cpu().numpy()[sl] if (eta and t.eta is not None) else None,
        dt_in_ms=ex.job.dt_ms, max_aperture_m=cfg.max_aperture_m, min_aperture_m=cfg.min_aperture_m,
        max_dip_deg=cfg.max_dip_degrees, taper_fraction=cfg.taper_fraction,
        apply_spreading=cfg.apply_spreading, apply_obliquity=cfg.apply_obliquity, grid_dx=cfg.grid_dx,
        grid_dy=cfg.grid_dy, traveltime=traveltime,
        aa_mode=_abi.AA_TRIANGLE if cfg.antialias == "triangle" else _abi.AA_NONE,
        aa_max_half_width=cfg.aa_max_half_width, eta_constant=cfg.vti_eta_constant)
    t0 = time.perf_counter()
    ref = orc.migrate(inp, fast=False)
    got = t.image.cpu().numpy()[sl].astype(np.float64)
    rel = float(np.linalg.norm(got - ref.image) / np.linalg.norm(ref.image))
    say(f"  oracle on {n_pillars} pillars x {len(rows)} traces in "
        f"{time.perf_counter() - t0:.1f} s: image rel-L2 {rel:.3e}")
    np.testing.assert_array_equal(ref.fold2d, t.fold.cpu().numpy()[sl])
    return rel


def _run_plan(ex, tiles=None, **kw):
    import torch

    ex.reset()
    plan = ex.plan(tiles=tiles, **kw)
    t0 = time.perf_counter()
    plan.run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.destroy()
    return dt


def test_reflectors_full_size(kernel):
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import reflector_job

    say("reflectors: building 200000 x 1501")
    job = reflector_job(device="cuda")
    ex = TileExecutor(kernel, job)
    torch.cuda.synchronize()
    dt = _run_plan(ex)
    say(f"reflectors: one plan over {len(ex.tiles)} tiles in {dt:.1f} s, C = {ex.contributions():.3e}")
    tiles = list(range(len(ex.tiles)))
    _fold_check(ex, tiles)
    say("reflectors: fold equal on all 65536 pillars")
    busiest = max(tiles, key=lambda i: int(ex.tiles[i].rows.shape[0]))
    rel_aa = _image_sample(ex, busiest, 16, _abi.TT_STRAIGHT)
    # the same pillars without anti-aliasing (the plain kernel)
    from dataclasses import replace

    ex.config = replace(ex.config, antialias="none")
    dt = _run_plan(ex, tiles=[busiest])
    rel_plain = _image_sample(ex, busiest, 16, _abi.TT_STRAIGHT)
    say(f"reflectors: busiest tile without AA in {dt:.1f} s; image rel-L2 AA {rel_aa:.3e}, plain {rel_plain:.3e}")
    # f32 sample positions carry ~1e-7 u samples of error (u up to 1500 here),
    # which on 30 Hz events is ~1e-5 of the image: the bar at this trace length
    assert rel_aa <= F32_FLOOR_1501 and rel_plain <= F32_FLOOR_1501


def test_bins_full_size(kernel):
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import bins_job

    say("bins: building 40 bins, 1.6M x 1501")
    job = bins_job(device="cuda")
    assert job.n_traces > 1_500_000 and len(job.meta["bin_start"]) == 41
    ex = TileExecutor(kernel, job)
    torch.cuda.synchronize()
    dt = _run_plan(ex)
    say(f"bins: one graph-replayed launch over {len(ex.tiles)} tiles x 40 bins in {dt:.1f} s, "
        f"C = {ex.contributions():.3e}")
    tiles = list(range(len(ex.tiles)))
    _fold_check(ex, tiles)
    say("bins: fold equal on all 65536 pillars")
    busiest = max(tiles, key=lambda i: int(ex.tiles[i].rows.shape[0]))
    rel = _image_sample(ex, busiest, 4, _abi.TT_STRAIGHT)
    assert rel <= F32_FLOOR_1501


def test_vti_full_size_region(kernel):
    import torch

    from pm_amd.autotune import region_tiles
    from pm_amd.lattice import SPARSE
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate, make_training_set
    from pm_amd.synthetic import vti_diffractor_job

    say("vti: building 500000 x 2001, 512 x 512 x 1500")
    job = vti_diffractor_job(device="cuda", tile=32)
    ex = TileExecutor(kernel, job)
    torch.cuda.synchronize()
    assert ex.shape == (512, 512, 1500) and ex.image_storage.numel() == 512 * 512 * 1500
    sub = region_tiles(ex, 224, 256, 224, 256)
    assert len(sub) == 1
    dt = _run_plan(ex, tiles=sub)
    say(f"vti: exact traveltime over the central 32 x 32 pillars in {dt:.1f} s")
    _fold_check(ex, sub)
    rel = _image_sample(ex, sub[0], 16, _abi.TT_VTI, eta=True)
    assert rel <= F32_FLOOR_2001
    exact = ex.tiles[sub[0]].image.double().clone()
    d_max = job.config.max_aperture_m + 1500.0
    norm = Normalizer(tau_max=1.5, d_max=d_max, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
    tr = make_training_set(job.cube, job.meta["eta_cube"], 10_000_000, d_max, norm, seed=1, t_max_ms=3000.0)
    sur = TraveltimeSurrogate.create(norm, seed=0)
    sur.weight_terms = 1
    t0 = time.perf_counter()
    sur.train(tr, steps=100_000, batch=65536, qat_fraction=0.25)  # as the benchmark trains it
    torch.cuda.synchronize()
    say(f"vti: surrogate trained in {time.perf_counter() - t0:.1f} s")
    del tr
    errs = ex.check_surrogate(sur, node_spacing=SPARSE, tiles=sub)
    assert errs[0] <= 0.25
    dt = _run_plan(ex, tiles=sub, traveltime=_abi.TT_VTI_LEARNED, surrogate=sur, node_spacing=SPARSE)
    learned = ex.tiles[sub[0]].image.double()
    rl = float(torch.linalg.norm(learned - exact) / torch.linalg.norm(exact))
    say(f"vti: learned traveltime in {dt:.1f} s, held-out {errs[0]:.3f} sample, image rel-L2 vs exact {rl:.3e}")
    assert rl <= 1e-3


def test_streamed_vti_bins_full_size(kernel, tmp_path):
    """The largest single-GPU job: 60 common-offset bins with lognormal counts
    (~2.4M traces x 2001 samples, 19.2 GB) written as a Zarr v2 array + a
    Parquet header table, evicted from the page cache, then streamed to HBM
    through pinned buffers while the bins migrate one by one, as they arrive,
    into a 1024 x 1024 x 2000 image (8.4 GB f32, the whole image and fold
    allocated) with the learned VTI traveltime and its per-tile check, on
    the stated central 32 x 32 pillars.  Fold equal on every pillar of the
    region to the oracle's over all traces; the learned image within relative
    L2 1e-3 of the exact-traveltime migration of the streamed traces."""
    import os

    import torch

    from pm_amd.io.zarr_v2 import write_headers, write_zarr_v2
    from pm_amd.lattice import SPARSE
    from pm_amd.pipeline.streaming import migrate_streamed
    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate, make_training_set
    from pm_amd.synthetic import vti_bins_job

    say("streamed: building 60 bins, ~2.4M x 2001, 1024 x 1024 x 2000")
    t0 = time.perf_counter()
    job = vti_bins_job(device="cuda")
    torch.cuda.synchronize()
    n, ns = job.n_traces, job.n_samples
    assert job.meta["bin_start"][-1] == n and 2_300_000 < n < 2_500_000 and ns == 2001
    say(f"streamed: {n} traces built in {time.perf_counter() - t0:.1f} s")
    bins = np.repeat(np.arange(60), np.diff(job.meta["bin_start"]))
    t0 = time.perf_counter()
    arr = write_zarr_v2(tmp_path / "traces.zarr", job.amplitudes, chunk_rows=10_000,
                        attrs={"sample_rate_ms": job.dt_ms, "start_time_ms": job.t_start_ms})
    write_headers(tmp_path / "headers.parquet", job.source_x, job.source_y, job.receiver_x, job.receiver_y,
                  extra={"BIN": bins})
    os.sync()
    for i in range(arr.n_chunks):   # the reads below come from the disk, not the page cache
        fd = os.open(tmp_path / "traces.zarr" / f"{i}.0", os.O_RDONLY)
        os.posix_fadvise(fd, 0, 0, os.POSIX_FADV_DONTNEED)
        os.close(fd)
    say(f"streamed: {n * ns * 4 / 1e9:.1f} GB written in {time.perf_counter() - t0:.1f} s")
    d_max = job.config.max_aperture_m + 1500.0
    norm = Normalizer(tau_max=2.0, d_max=d_max, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
    tr = make_training_set(job.cube, job.meta["eta_cube"], 10_000_000, d_max, norm, seed=1, t_max_ms=4000.0)
    sur = TraveltimeSurrogate.create(norm, seed=0)
    t0 = time.perf_counter()
    sur.train(tr, steps=100_000, batch=65536, qat_fraction=0.25)
    torch.cuda.synchronize()
    say(f"streamed: surrogate trained in {time.perf_counter() - t0:.1f} s")
    del tr
    x_axis, y_axis, t_axis, cfg, cube, ecube = (job.x_axis, job.y_axis, job.t_axis_ms, job.config, job.cube,
                                                job.meta["eta_cube"])
    mx, my = job.midpoint_x, job.midpoint_y
    del job
    torch.cuda.empty_cache()
    region = (496, 528, 496, 528)
    res = migrate_streamed(kernel, tmp_path / "traces.zarr", tmp_path / "headers.parquet", x_axis, y_axis,
                           t_axis, cfg, cube=cube, eta_cube=ecube, tile=16, traveltime=_abi.TT_VTI_LEARNED,
                           surrogate=sur, node_spacing=SPARSE, region=region, n_buffers=8, n_threads=4)
    ex = res.executor
    sub = sorted((i for i, t in enumerate(ex.tiles) if t.x0 >= 496 and t.x1 <= 528 and t.y0 >= 496 and t.y1 <= 528),
                 key=lambda i: (ex.tiles[i].x0, ex.tiles[i].y0))
    assert len(sub) == 4
    errs = [ex.tiles[i].heldout_error for i in sub]
    say(f"streamed: {res.groups} bins, {res.stats.chunks} chunks, {res.stats.bytes_read / 1e9:.1f} GB in "
        f"{res.wall_s:.1f} s end to end ({res.stats.bytes_read / 1e9 / res.wall_s:.2f} GB/s), held-out "
        f"{[round(e, 3) for e in errs]}, errors {res.stats.errors}")
    assert res.groups == 60 and res.stats.chunks == arr.n_chunks and not res.stats.errors
    assert ex.shape == (1024, 1024, 2000) and ex.image_storage.numel() == 1024 * 1024 * 2000
    assert all(ex.tiles[i].learned_ok for i in sub)
    learned = torch.cat([torch.cat([ex.tiles[i].image.double() for i in sub[r:r + 2]], 1)
                         for r in range(0, 4, 2)], 0)
    fold = torch.cat([torch.cat([ex.tiles[i].fold for i in sub[r:r + 2]], 1) for r in range(0, 4, 2)], 0)
    want = orc.fold(x_axis[496:528], y_axis[496:528], mx, my, cfg.max_aperture_m)
    np.testing.assert_array_equal(fold.cpu().numpy(), want)
    t0 = time.perf_counter()
    _run_plan(ex, tiles=sub)   # exact traveltime, the streamed traces
    exact = torch.cat([torch.cat([ex.tiles[i].image.double() for i in sub[r:r + 2]], 1)
                       for r in range(0, 4, 2)], 0)
    rl = float(torch.linalg.norm(learned - exact) / torch.linalg.norm(exact))
    say(f"streamed: exact pass {time.perf_counter() - t0:.1f} s; learned image rel-L2 vs exact {rl:.3e}")
    assert rl <= 1e-3
