"""Benchmark: Gcontrib/s of the MI355X PSTM backend on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload reflectors]

A step is one migration pass over the whole synthetic job with every input
already resident in HBM: for the default workload, 16 output tiles (64 x 64
pillars x 1000 samples each) over 200,000 traces of 1501 samples with 3-D
v_rms and triangle anti-aliasing, run as one plan (one persistent gather
launch fed by a device work queue, replayed from a captured hipGraph).  A
contribution is one (trace, output sample) pair inside the maximum aperture,
C = sum of the 3-D fold; the rate is C / step time.  The roofline entry uses E
(contributions that pass the time-dependent aperture and the sample bounds) x
the frozen FLOPs per contribution over the gather launch's own duration,
measured with HIP events on its stream.  The CPU baseline is the C++ oracle
(built with -O3 -ffast-math, OpenMP) on a bounded pillar sample of the same
job.  The learned_vs_exact entry times VTI migration with the learned and the
exact traveltime on a reduced VTI diffractor job.

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("Gcontrib/s (trace×output-sample), learned-TT vs exact-TT, image "
          "relL2≤1e-3")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_job(name: str, device: str):
    from pm_amd.synthetic import point_diffractor_job, reflector_job

    if name == "reflectors":
        return reflector_job(device=device), (
            "reflectors: 200000 traces x 1501 samples (2 ms), 40 dipping planes + 200 "
            "diffractors, 3-D v_rms cube 64x64x151, output 256x256 bins at 12.5 m x 1000 "
            "samples in 16 tiles of 64x64, aperture 2500 m, dip 60, triangle AA, f32")
    if name == "bins":
        from pm_amd.synthetic import bins_job

        return bins_job(device=device), (
            "bins: 40 common-offset bins (50 m spacing), lognormal counts (median 30k, 5k-120k), "
            "1.6M traces x 1501 samples, reflector model, output 256x256 x 1000 samples in 16 tiles "
            "of 64x64, one shared image, no AA, f32")
    if name == "point_diffractor":
        return point_diffractor_job(device=device), (
            "point diffractor: 4096 traces x 1001 samples, output 64x64 x 500, aperture "
            "1000 m, weights on, no AA")
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(ex, job, budget_s: float) -> dict:
    """Oracle (-O3 -ffast-math, OpenMP) on a pillar sample of the busiest
    tile, all of that tile's traces, all output samples."""
    import numpy as np

    from oracle import oracle as orc
    from pm_amd.kernels import _abi

    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    tile = max(ex.tiles, key=lambda t: int(t.rows.shape[0]))
    rows = tile.rows.cpu().numpy().astype(np.int64)
    h = ex.host_traces(rows)
    cfg = ex.config
    ix = (tile.x1 - tile.x0) // 2
    v3 = tile.vrms.cpu().numpy() if tile.vrms.dim() == 3 else None
    v1 = tile.vrms.cpu().numpy() if tile.vrms.dim() == 1 else None
    xa, ya = tile.x_axis.cpu().numpy(), tile.y_axis.cpu().numpy()

    def run(n_p: int, n_x: int = 1):
        inp = orc.MigrateInputs(
            amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
            receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
            midpoint_y=h["midpoint_y"], x_axis=xa[ix:ix + n_x], y_axis=ya[:n_p],
            t_axis_ms=job.t_axis_ms, vrms=(v3[ix:ix + n_x, :n_p] if v3 is not None else v1),
            dt_in_ms=job.dt_ms, max_aperture_m=cfg.max_aperture_m,
            min_aperture_m=cfg.min_aperture_m, max_dip_deg=cfg.max_dip_degrees,
            taper_fraction=cfg.taper_fraction, apply_spreading=cfg.apply_spreading,
            apply_obliquity=cfg.apply_obliquity, grid_dx=cfg.grid_dx, grid_dy=cfg.grid_dy,
            aa_mode=_abi.AA_TRIANGLE if cfg.antialias == "triangle" else _abi.AA_NONE,
            aa_max_half_width=cfg.aa_max_half_width)
        t0 = time.perf_counter()
        res = orc.migrate(inp, threads=threads, fast=True)
        return time.perf_counter() - t0, int(res.fold2d.sum()) * len(job.t_axis_ms), res

    ny = tile.y1 - tile.y0
    run(1)                      # warm-up: library load, OpenMP thread start
    dt1, _, _ = run(ny)         # one inline: every thread busy
    n_x = int(max(1, min(tile.x1 - tile.x0 - ix, round(budget_s / max(dt1, 1e-3)))))
    dt, c, _ = run(ny, n_x)
    log(f"cpu baseline: {n_x} x {ny} pillars x {len(rows)} traces in {dt:.1f} s on {threads} threads")
    return {"value": c / dt / 1e9, "unit": "Gcontrib/s", "cores": threads, "kind": "port",
            "sample": (f"{n_x} x {ny} pillars (whole inlines of the busiest tile) x all {len(rows)} "
                       f"traces of that tile x {len(job.t_axis_ms)} samples, same settings; "
                       f"C_sample={c}, {dt:.2f} s; oracle C++ FP64 -O3 -ffast-math OpenMP")}


# HBM bytes per gather launch from rocprofv3 FETCH_SIZE / WRITE_SIZE runs (KiB per
# dispatch scaled by true bytes / counter of the calibration kernel), per workload
# (the committed rocprofv3 summary of the reflector gather: FETCH x 2.000 + WRITE
# x 1.000, one plan launch)
TRAFFIC: dict = {"reflectors": (8_834_774_528 * 2.000 + 256_297 * 1.000) * 1024}


def learned_vs_exact(kernel, args) -> dict:
    """Config-4 family, reduced to fit a bench run: VTI diffractor job, the
    surrogate trained on the job's own v_rms / eta model by the library's
    kernels, per-tile held-out check, then one pass of every tile with the
    exact VTI traveltime and one with the learned traveltime (each a plan:
    one persistent graph launch), timed with HIP events; image error of the
    learned image against the exact one, and of the isotropic image (the
    method without the network) against it."""
    import torch
    from dataclasses import replace

    from pm_amd.kernels import _abi
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate, make_training_set
    from pm_amd.synthetic import vti_diffractor_job

    t0 = time.perf_counter()
    job = vti_diffractor_job(n_traces=args.vti_traces, n_samples=2001, n_out=args.vti_out, nt=1500, tile=64,
                             device="cuda")
    ex = TileExecutor(kernel, job)
    vc, ec = job.cube, job.meta["eta_cube"]
    d_max = job.config.max_aperture_m + 1500.0
    norm = Normalizer(tau_max=1.5, d_max=d_max, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
    torch.cuda.synchronize()
    log(f"vti job ready in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    tr = make_training_set(vc, ec, args.train_samples, d_max, norm, seed=1, t_max_ms=3000.0)
    sur = TraveltimeSurrogate.create(norm, seed=0).train(tr, steps=args.train_steps, batch=args.train_batch,
                                                          qat_fraction=0.25, log_every=5000, verbose=True)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del tr
    log(f"surrogate trained in {train_s:.1f} s: {sur.history}")
    K = args.node_spacing
    errs = ex.check_surrogate(sur, node_spacing=K)
    learned_tiles = [i for i, t in enumerate(ex.tiles) if t.learned_ok]
    exact_tiles = [i for i, t in enumerate(ex.tiles) if not t.learned_ok]
    lib = _abi.load_library()

    def timed(plan_list):
        ex.reset()
        for p in plan_list:
            p.run(use_graph=True)       # capture + warm
        torch.cuda.synchronize()
        ex.reset()
        torch.cuda.synchronize()
        lib.pstm_hip_profile_reset(1)
        for p in plan_list:
            p.run(use_graph=True)
        kms, nl = ctypes.c_double(), ctypes.c_longlong()
        lib.pstm_hip_profile_read(ctypes.byref(kms), ctypes.byref(nl))
        lib.pstm_hip_profile_reset(0)
        img, _ = ex.assemble()
        torch.cuda.synchronize()
        return kms.value, img.double()

    from pm_amd.autotune import autotune

    vt = autotune(kernel, ex, candidates=[(2, 10, 0), (2, 16, 1), (4, 8, 0), (4, 16, 1)])
    log(f"vti autotune: mw={vt.mw} batch={vt.batch} pipe={vt.pipe}")
    exact_plan = ex.plan()
    ms_exact, img_exact = timed([exact_plan])
    plans = []
    if learned_tiles:
        plans.append(ex.plan(tiles=learned_tiles, traveltime=_abi.TT_VTI_LEARNED, surrogate=sur, node_spacing=K))
    if exact_tiles:
        plans.append(ex.plan(tiles=exact_tiles))
    ms_learned, img_learned = timed(plans)
    c_total = ex.contributions()
    saved = ex.config
    ex.config = replace(saved, kernel_type="straight_ray", vti_enabled=False)
    iso_plan = ex.plan()
    _, img_iso = timed([iso_plan])
    ex.config = saved
    nrm = float(torch.linalg.norm(img_exact))
    rel = float(torch.linalg.norm(img_learned - img_exact)) / nrm
    rel_iso = float(torch.linalg.norm(img_iso - img_exact)) / nrm
    for p in [exact_plan, iso_plan, *plans]:
        p.destroy()
    out = {
        "workload": (f"VTI diffractors: {job.n_traces} traces x 2001 samples, 3-D v_rms 1500-4500 and eta "
                     f"0-0.2 on 128x128x201 grids, output {args.vti_out}x{args.vti_out} x 1500 samples in "
                     f"{len(ex.tiles)} tiles, aperture 2500 m"),
        "surrogate": {"train_samples": args.train_samples, "steps": args.train_steps, "batch": args.train_batch,
                      "train_s": round(train_s, 1), "node_spacing": K,
                      "heldout_max_samples_per_tile": [round(e, 4) for e in errs],
                      "tiles_learned": len(learned_tiles), "tiles_exact_fallback": len(exact_tiles)},
        "contributions": c_total,
        "exact_gcontrib_s": round(c_total / (ms_exact / 1e3) / 1e9, 3),
        "learned_gcontrib_s": round(c_total / (ms_learned / 1e3) / 1e9, 3),
        "exact_ms": round(ms_exact, 2), "learned_ms": round(ms_learned, 2),
        "launch_shape": {"mw": vt.mw, "batch": vt.batch, "pipe": vt.pipe},
        "learned_speedup": round(ms_exact / ms_learned, 3),
        "image_rel_l2_learned_vs_exact": rel,
        "image_rel_l2_without_network_vs_exact": rel_iso,
    }
    log(f"learned vs exact: {out}")
    return out


def streamed_section(kernel, args) -> dict:
    """Config-5 family, scaled: common-offset bins with lognormal counts,
    2001-sample traces, written to local disk as an uncompressed Zarr v2 array
    (1000-trace chunks) + Parquet headers, then migrated while streaming
    (pinned buffers, copy stream, one plan per bin launched as the bin
    arrives) and, for comparison, from memory (one plan)."""
    import shutil
    import tempfile

    import numpy as np
    import torch

    from pm_amd.io.zarr_v2 import write_headers, write_zarr_v2
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.pipeline.streaming import migrate_streamed
    from pm_amd.synthetic import bins_job

    job = bins_job(n_bins=args.stream_bins, total=args.stream_traces, n_samples=2001, n_out=256, nt=1000,
                   tile=64, device="cuda")
    d = Path(tempfile.mkdtemp(prefix="pstm_stream_", dir=os.environ.get("TMPDIR", "/tmp")))
    try:
        t0 = time.perf_counter()
        bins = np.repeat(np.arange(args.stream_bins), np.diff(job.meta["bin_start"]))
        write_zarr_v2(d / "traces.zarr", job.amplitudes, chunk_rows=1000,
                      attrs={"sample_rate_ms": job.dt_ms, "start_time_ms": job.t_start_ms})
        write_headers(d / "headers.parquet", job.source_x, job.source_y, job.receiver_x, job.receiver_y,
                      extra={"BIN": bins})
        write_s = time.perf_counter() - t0
        nbytes = job.n_traces * job.n_samples * 4
        log(f"stream: wrote {nbytes / 1e9:.2f} GB in {write_s:.1f} s")
        # in memory: one plan over every tile (file order, as streamed)
        ex = TileExecutor(kernel, job, spatial_sort=False)
        plan = ex.plan()
        ex.reset()
        plan.run()                      # capture
        ex.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan.run()
        torch.cuda.synchronize()
        mem_s = time.perf_counter() - t0
        img_mem, fold_mem = ex.assemble()
        img_mem = img_mem.clone()
        plan.destroy()
        del ex
        torch.cuda.empty_cache()
        res = migrate_streamed(kernel, d / "traces.zarr", d / "headers.parquet", job.x_axis, job.y_axis,
                               job.t_axis_ms, job.config, cube=job.cube, tile=job.tile, n_buffers=16,
                               n_threads=8)
        rel = float(torch.linalg.norm((res.image - img_mem).double()) / torch.linalg.norm(img_mem.double()))
        out = {"workload": (f"{args.stream_bins} common-offset bins, {job.n_traces} traces x 2001 samples "
                            f"({nbytes / 1e9:.2f} GB f32) in uncompressed Zarr v2, 256x256 x 1000 output"),
               "bytes": nbytes, "groups": res.groups, "chunks": res.stats.chunks,
               "streamed_wall_s": round(res.wall_s, 2), "in_memory_gather_s": round(mem_s, 2),
               "streaming_overhead_s": round(res.wall_s - mem_s, 2),
               "read_threads_busy_s": round(res.stats.read_s, 2),
               "file_read_gb_s": round(nbytes / 1e9 / max(res.stats.read_s, 1e-9), 2),
               "note": "files written just before: reads are served from the page cache",
               "fold_equal": bool(torch.equal(res.fold, fold_mem)), "image_rel_l2_vs_in_memory": rel}
        log(f"stream: {out}")
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def output_section(kernel, ex) -> dict:
    """The output step on the migrated job: per-bin header sums over every
    tile's trace query (HIP atomics), the gather-mode fold normalisation of
    the whole volume (HBM-bound), and the stack finalize write-out
    (migrated_stack.zarr + fold.zarr + bin_headers.parquet) into a temporary
    directory that is removed afterwards."""
    import shutil
    import tempfile

    import torch

    from pm_amd.pipeline.output import HeaderSums

    def timed_ms(fn, reps):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    nx, ny, nt = ex.shape
    rows = sum(int(t.rows.shape[0]) for t in ex.tiles)
    sums = HeaderSums.zeros(nx, ny, ex.device)

    def headers():
        for s in (sums.trace_count, sums.offset_sum, sums.azimuth_sin_sum, sums.azimuth_cos_sum):
            s.zero_()
        ex.header_sums(sums)

    h_ms = timed_ms(headers, 5)
    image, fold = ex.assemble()
    out = torch.empty_like(image)
    n_ms = timed_ms(lambda: kernel.fold_normalize(image, fold, out), 20)
    n_bytes = 2 * image.numel() * 4 + fold.numel() * 4
    d = tempfile.mkdtemp(prefix="pm_amd_out_")
    try:
        torch.cuda.synchronize()
        st = ex.finalize(d, headers=sums)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    log(f"output: headers {h_ms:.2f} ms, normalise {n_ms:.3f} ms, write {st.bytes / 1e9:.2f} GB "
        f"in {st.seconds:.2f} s")
    return {"header_sums_ms": round(h_ms, 3), "header_rows": rows,
            "fold_normalize_ms": round(n_ms, 4), "fold_normalize_GBps": round(n_bytes / n_ms / 1e6, 1),
            "fold_normalize_frac_of_hbm": round(n_bytes / n_ms / 1e6 / 8000.0, 3),
            "write_GB": round(st.bytes / 1e9, 3), "write_s": round(st.seconds, 3),
            "write_GBps": round(st.gb_per_s, 2),
            "files": ["migrated_stack.zarr", "fold.zarr", "bin_headers.parquet"]}


def time_variant_section(kernel, job, uniform_image, uniform_ms: float, uniform_e: int) -> dict:
    """The same job migrated on the time-variant windows of the reference's
    default frequency table (80 Hz at 0 s to 20 Hz at 5 s): launch shape
    autotuned for the window axis, E counted, then one plan timed with events
    around one graph replay after the capturing run; the window image
    resampled onto the uniform axis on the device and compared with the
    uniform image of the main measurement."""
    import torch

    from pm_amd import time_variant as tv
    from pm_amd.autotune import autotune
    from pm_amd.pipeline.executor import TileExecutor

    t = job.t_axis_ms
    table = tv.FrequencyTimeTable.from_pairs(tv.CONFIG_DEFAULT_TABLE)
    w = tv.compute_time_windows(float(t[0]), float(t[-1]), float(t[1] - t[0]), table)
    ex = TileExecutor(kernel, job, time_windows=w)
    tune = autotune(kernel, ex, candidates=[(2, 8, 0), (2, 16, 1), (1, 8, 0), (1, 16, 1)])
    cplan = ex.plan(count_effective=True)
    cplan.run(use_graph=False)
    torch.cuda.synchronize()
    e_tv = int(cplan.counter.item())
    cplan.destroy()
    plan = ex.plan()
    ex.reset()
    plan.run()
    torch.cuda.synchronize()
    ex.reset()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    plan.run()
    b.record()
    b.synchronize()
    tv_ms = a.elapsed_time(b)
    plan.destroy()
    img, _ = ex.assemble()
    a.record()
    full = ex.assemble_uniform()
    b.record()
    b.synchronize()
    rs_ms = a.elapsed_time(b)
    ref = uniform_image.double()
    rel = float((full.double() - ref).norm() / ref.norm())
    n_tv = int(ex.nt)
    nbytes = img.numel() * 4 + full.numel() * 4
    log(f"time-variant: {len(w)} windows, {n_tv} of {len(t)} samples, {tv_ms:.0f} ms vs {uniform_ms:.0f} ms, "
        f"rel-L2 {rel:.3e}")
    del ex, img, full
    torch.cuda.empty_cache()
    return {"table": [list(p) for p in tv.CONFIG_DEFAULT_TABLE], "windows": len(w),
            "factors": [x.downsample_factor for x in w], "samples": n_tv, "uniform_samples": len(t),
            "migrate_ms": round(tv_ms, 1), "uniform_migrate_ms": round(uniform_ms, 1),
            "speedup": round(uniform_ms / tv_ms, 3), "effective_contributions": e_tv,
            "effective_fraction_of_uniform": round(e_tv / uniform_e, 4),
            "launch_shape": {"mw": tune.mw, "batch": tune.batch, "pipe": tune.pipe},
            "resample_ms": round(rs_ms, 3), "resample_GBps_incl_assemble": round(nbytes / rs_ms / 1e6, 1),
            "rel_l2_vs_uniform_image": round(rel, 5)}


def interpolation_section(ex, methods=("linear", "cubic", "sinc8", "sinc16", "lanczos3")) -> dict:
    """Cost of the trace interpolators: the job without anti-aliasing (the
    interpolators other than linear need it off), the busiest tile as a plan
    per method, one eager run timed with events (each method logged as it
    ends)."""
    from dataclasses import replace

    import torch

    from pm_amd.kernels import _abi

    lib = _abi.load_library()
    lib.pstm_hip_set_tuning(0, 0)       # each method at its built-in launch shape
    lib.pstm_hip_set_pipeline(-1)
    busiest = [max(range(len(ex.tiles)), key=lambda i: int(ex.tiles[i].rows.shape[0]))]
    saved = ex.config
    out = {}
    try:
        for m in methods:
            ex.config = replace(saved, antialias="none", interpolation_method=m)
            plan = ex.plan(tiles=busiest)
            ex.reset()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            plan.run(use_graph=False)
            b.record()
            b.synchronize()
            out[m] = round(a.elapsed_time(b), 1)
            plan.destroy()
            log(f"interpolation {m}: {out[m]} ms (busiest tile, no AA)")
    finally:
        ex.config = saved
        ex.reset()
    return {"ms_busiest_tile_no_aa": out,
            "relative_to_linear": {m: round(v / out["linear"], 3) for m, v in out.items()}}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="reflectors", choices=["reflectors", "point_diffractor", "bins"])
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip timing candidate launch shapes on the busiest tile")
    ap.add_argument("--candidates", default="4x8x0,4x16x1,2x10x0,4x3x1",
                    help="autotune candidates as MWxBATCHxPIPE,... (PIPE: 0 off, 1 pipelined staging)")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: one elementwise kernel over 256 MiB (traffic calibration), "
                         "then the steps as eager launches; no counting pass, no CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample (0 = skip)")
    ap.add_argument("--no-learned", dest="learned", action="store_false",
                    help="skip the learned-vs-exact VTI traveltime comparison")
    ap.add_argument("--vti-traces", type=int, default=100_000)
    ap.add_argument("--vti-out", type=int, default=128)
    ap.add_argument("--train-samples", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=80_000)
    ap.add_argument("--train-batch", type=int, default=16_384)
    ap.add_argument("--node-spacing", type=int, default=16)
    ap.add_argument("--no-interpolation", dest="interpolation", action="store_false",
                    help="skip the interpolator timing section")
    ap.add_argument("--no-time-variant", dest="time_variant", action="store_false",
                    help="skip the time-variant sampling section")
    ap.add_argument("--no-output", dest="output", action="store_false",
                    help="skip the output-step section (header sums, normalisation, finalize write-out)")
    ap.add_argument("--stream", action="store_true",
                    help="also measure migration streamed from disk (scaled config-5 bins job)")
    ap.add_argument("--stream-bins", type=int, default=60)
    ap.add_argument("--stream-traces", type=int, default=240_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the migration shards "
              "by output tile with no collective, so N GPUs run N independent tile sets",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from pm_amd import roofline
    from pm_amd.kernels import HipKernel
    from pm_amd.kernels import _abi
    from pm_amd.pipeline.executor import TileExecutor

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = "cuda"
    t0 = time.perf_counter()
    job, desc = make_job(args.workload, dev)
    torch.cuda.synchronize()
    log(f"job built in {time.perf_counter() - t0:.1f} s: {job.n_traces} traces x {job.n_samples}")
    kernel = HipKernel()
    kernel.initialize(job.config)
    ex = TileExecutor(kernel, job)
    torch.cuda.synchronize()
    log(f"executor ready: {len(ex.tiles)} tiles, rows/tile {[int(t.rows.shape[0]) for t in ex.tiles][:4]}...")
    tune = None
    if args.autotune and not args.profile:
        from pm_amd.autotune import autotune

        cands = None
        if args.candidates:
            cands = [tuple(int(v) for v in c.split("x")) for c in args.candidates.split(",")]
        tune = autotune(kernel, ex, candidates=cands)
        log(f"autotune: mw={tune.mw} batch={tune.batch} pipe={tune.pipe} ({tune.ms:.1f} ms on the 4 busiest tiles)")
    lib = _abi.load_library()

    if args.profile:
        # elementwise kernel of known size (a name nothing else here launches):
        # reference point for the HBM byte counters
        a = torch.full((64 * 2**20,), 7, dtype=torch.int32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.bitwise_not(a, out=b)
        torch.cuda.synchronize()
        log("calibration bitwise_not: 268435456 bytes read, 268435456 bytes written")
        del a, b
        plan = ex.plan()
        for _ in range(args.steps):
            ex.reset()
            plan.run(use_graph=False)
            torch.cuda.synchronize()
            log("profile step done")
        plan.destroy()
        kernel.cleanup()
        return

    # E (counting plan, not timed) and C
    cplan = ex.plan(count_effective=True)
    ex.reset()
    cplan.run(use_graph=False)
    torch.cuda.synchronize()
    e_total = int(cplan.counter.item())
    cplan.destroy()
    c_total = ex.contributions()
    img, _ = ex.assemble()
    assert bool(torch.isfinite(img).all()) and float(img.abs().max()) > 0, "empty or non-finite image"
    log(f"C = {c_total:.4e} contributions, E = {e_total:.4e}")

    plan = ex.plan()
    for _ in range(args.warmup):
        ex.reset()
        plan.run()           # the first run captures the graph
    torch.cuda.synchronize()
    # HIP events on the launch stream around every graph launch (the gather
    # plus a few small kernels: fold add, queue reset, AA row integration)
    lib.pstm_hip_profile_reset(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ex.reset()
        plan.run()
        torch.cuda.synchronize()
        log("step done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1000.0 / args.steps
    value = c_total / (elapsed / args.steps) / 1e9
    kms, nl = ctypes.c_double(), ctypes.c_longlong()
    lib.pstm_hip_profile_read(ctypes.byref(kms), ctypes.byref(nl))
    lib.pstm_hip_profile_reset(0)
    launch = plan.info()
    plan.destroy()

    cfg = job.config
    f = roofline.flops_per_contribution(cfg.kernel_type, cfg.antialias,
                                        bool(cfg.apply_spreading or cfg.apply_obliquity))
    launches = int(nl.value)
    launch_ms = kms.value / max(launches, 1)
    e_per_launch = e_total          # one gather launch covers every tile
    achieved = f * e_per_launch / (launch_ms / 1000.0) / 1e12
    peak = roofline.MI355X_FP32_VALU_TFLOPS
    rl = {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
          "frac": round(achieved / peak, 4), "traffic": TRAFFIC.get(args.workload),
          "kernel": "k_migrate (gather, all tiles in one persistent launch; timed as its graph)",
          "flops_per_contribution": f,
          "effective_contributions_per_launch": int(e_per_launch),
          "avg_launch_ms": round(launch_ms, 3), "launches_timed": launches,
          "frac_of_unpacked_fp32": round(achieved / roofline.MI355X_FP32_VALU_UNPACKED_TFLOPS, 4)}
    output = output_section(kernel, ex) if args.output else None
    tvar = None
    if args.time_variant:
        img_u, _ = ex.assemble()
        tvar = time_variant_section(kernel, job, img_u, ms_step, e_total)
        del img_u
    interp = interpolation_section(ex) if args.interpolation else None
    # the VTI job has its own launch shape: back to the built-in choice
    lib.pstm_hip_set_tuning(0, 0)
    lib.pstm_hip_set_pipeline(-1)
    learned = learned_vs_exact(kernel, args) if args.learned else None
    streamed = streamed_section(kernel, args) if args.stream else None
    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(ex, job, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcontrib/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": desc, "traces": job.n_traces, "samples": job.n_samples,
                   "output": list(ex.shape), "tiles": len(ex.tiles),
                   "contributions": c_total, "effective_contributions": e_total,
                   "traveltime": "exact straight-ray DSR", "parallelism": "1 GPU",
                   "launch": "one persistent hipGraph-captured gather per step over every tile"},
        "roofline": rl, "cpu_baseline": cpu,
        "launch": launch,
    }
    if learned is not None:
        out["learned_vs_exact"] = learned
    if streamed is not None:
        out["streamed"] = streamed
    if output is not None:
        out["output"] = output
    if tvar is not None:
        out["time_variant"] = tvar
    if interp is not None:
        out["interpolation"] = interp
    if tune is not None:
        out["autotune"] = tune.as_dict()
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    kernel.cleanup()


if __name__ == "__main__":
    main()
