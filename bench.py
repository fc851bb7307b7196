"""Benchmark: Gcontrib/s of learned-traveltime VTI migration on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the learned-vs-exact traveltime job): 500,000 traces x 2001
samples (2 ms) with 300 random diffractors at their exact VTI times, 3-D
v_rms (1500-4500 m/s) and eta (0-0.2) fields on 128 x 128 x 201 grids, output
512 x 512 bins at 12.5 m x 1500 samples, aperture 2500 m, max dip 60 deg,
spreading and obliquity weights, straight-ray VTI traveltime.  Every input is
resident in HBM and the whole job is built (all traces, the full velocity and
eta volumes, the full image and fold, every tile's trace list), but a step is
one migration pass over a FIXED REGION of its output: the central
``--region`` x ``--region`` pillars (32 x 32 by default), in the tiles of the
edge the autotuner picks: a pass over all 512 x 512 pillars takes minutes,
and a run of 20 + 5 steps must fit in a few minutes.  C counts exactly
the region's contributions (sum of its 3-D fold), so the rate is that of the
region.  The autotuner times tile edges (output bins per tile) x launch
shapes (samples per wave, traces per batch, pipelining) with the exact
traveltime over twice the region and reports its table.

The traveltime surrogate (5-64-64-64-64-2 SiLU MLP) is trained by the
library's HIP kernels on 10M samples of the job's own v_rms / eta model
(training time reported, not part of a step), checked per tile against the
exact VTI time (0.25-sample held-out bar; failing tiles run the exact
traveltime), and a step is one persistent, hipGraph-captured gather launch
per plan (learned tiles, fallback tiles).  The same subset is migrated with
the exact VTI traveltime (``exact_ms``, ``learned_speedup``) and the learned
image is compared with the exact one (relative L2; target <= 1e-3).

Also reported: the roofline of the learned gather (VALU FLOPs per effective
contribution x E over the launch time, measured with HIP events on the
launch stream), the plain straight-ray gather on the same subset and on the
4096-trace point-diffractor job (both against the FP32 VALU peak), and the
CPU baseline (the FP64 oracle, -O3 -ffast-math, OpenMP, on a pillar sample of
the subset).  Side measurements on the reflector job (output step,
time-variant sampling, interpolators, the drop-in migrate_tile call with host
arrays, streaming) are opt-in: ``--sections``.

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


# False in profiler runs (--profile): every plan is launched directly, no
# hipGraph capture or replay, and the surrogate trains step by step
USE_GRAPH = True


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def central_region(nx: int, ny: int, region: int, align: int) -> tuple[int, int, int, int]:
    """The region x region pillar block nearest the grid centre whose corner
    sits on a multiple of ``align`` (so that tiles of every edge dividing
    ``align`` cover it exactly): (x0, x1, y0, y1)."""
    x0, y0 = (nx - region) // 2, (ny - region) // 2
    x0 -= x0 % align
    y0 -= y0 % align
    return x0, x0 + region, y0, y0 + region


def reset_tiles(ex, tiles) -> None:
    for i in tiles:
        ex.tiles[i].image.zero_()
        ex.tiles[i].fold.zero_()


def subset_image(ex, tiles):
    import torch

    return torch.cat([ex.tiles[i].image.reshape(-1) for i in tiles]).double()


def subset_contributions(ex, tiles) -> int:
    return int(sum(int(ex.tiles[i].fold.sum()) for i in tiles)) * ex.nt


def plan_launch_ms(lib, plans, reps: int = 1) -> float:
    """Average duration of one pass over the plans (summed launch times, HIP
    events on the launch stream, recorded by the library)."""
    import torch

    torch.cuda.synchronize()
    lib.pstm_hip_profile_reset(1)
    for _ in range(reps):
        for p in plans:
            p.run(use_graph=USE_GRAPH)
    kms, nl = ctypes.c_double(), ctypes.c_longlong()
    lib.pstm_hip_profile_read(ctypes.byref(kms), ctypes.byref(nl))
    lib.pstm_hip_profile_reset(0)
    return kms.value / reps


def calibration_kernel() -> None:
    """Elementwise kernel of known size (a name nothing else here launches):
    the reference point for the HBM byte counters of a profiler run."""
    import torch

    a = torch.full((64 * 2**20,), 7, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.bitwise_not(a, out=b)
    torch.cuda.synchronize()
    log("calibration bitwise_not: 268435456 bytes read, 268435456 bytes written")
    del a, b


def count_effective(ex, tiles, **kw) -> int:
    import torch

    reset_tiles(ex, tiles)
    p = ex.plan(tiles=tiles, count_effective=True, **kw)
    p.run(use_graph=False)
    torch.cuda.synchronize()
    e = int(p.counter.item())
    p.destroy()
    reset_tiles(ex, tiles)
    return e


def roofline_entry(f: int, e: int, launch_ms: float, kernel: str) -> dict:
    from pm_amd import roofline

    achieved = f * e / (launch_ms / 1000.0) / 1e12
    peak = roofline.MI355X_FP32_VALU_TFLOPS
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": None, "kernel": kernel,
            "flops_per_contribution": f, "effective_contributions_per_launch": int(e),
            "avg_launch_ms": round(launch_ms, 3),
            "frac_of_unpacked_fp32": round(achieved / roofline.MI355X_FP32_VALU_UNPACKED_TFLOPS, 4)}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ex, tile_index: int, budget_s: float, traveltime: int) -> dict:
    """The FP64 oracle (-O3 -ffast-math, OpenMP on every host core) on whole
    inlines of one tile: all of that tile's traces, all output samples, the
    same settings and traveltime; the inline count is sized to about
    budget_s of CPU time."""
    import numpy as np

    from oracle import oracle as orc
    from pm_amd.kernels import _abi

    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    tile = ex.tiles[tile_index]
    rows = tile.rows.cpu().numpy().astype(np.int64)
    h = ex.host_traces(rows)
    cfg = ex.config
    job = ex.job
    v3 = tile.vrms.cpu().numpy() if tile.vrms.dim() == 3 else None
    v1 = tile.vrms.cpu().numpy() if tile.vrms.dim() == 1 else None
    e3 = tile.eta.cpu().numpy() if tile.eta is not None else None
    xa, ya = tile.x_axis.cpu().numpy(), tile.y_axis.cpu().numpy()
    ny = tile.y1 - tile.y0
    ix0 = (tile.x1 - tile.x0) // 2

    def run(n_p: int, n_x: int = 1):
        sl = (slice(ix0, ix0 + n_x), slice(0, n_p))
        inp = orc.MigrateInputs(
            amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
            receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
            midpoint_y=h["midpoint_y"], x_axis=xa[sl[0]], y_axis=ya[sl[1]],
            t_axis_ms=np.asarray(job.t_axis_ms, float), vrms=(v3[sl] if v3 is not None else v1),
            eta=(e3[sl] if e3 is not None else None), dt_in_ms=job.dt_ms,
            max_aperture_m=cfg.max_aperture_m, min_aperture_m=cfg.min_aperture_m,
            max_dip_deg=cfg.max_dip_degrees, taper_fraction=cfg.taper_fraction,
            apply_spreading=cfg.apply_spreading, apply_obliquity=cfg.apply_obliquity,
            grid_dx=cfg.grid_dx, grid_dy=cfg.grid_dy, traveltime=traveltime,
            aa_mode=_abi.AA_TRIANGLE if cfg.antialias == "triangle" else _abi.AA_NONE,
            aa_max_half_width=cfg.aa_max_half_width, eta_constant=cfg.vti_eta_constant)
        t0 = time.perf_counter()
        res = orc.migrate(inp, threads=threads, fast=True)
        return time.perf_counter() - t0, int(res.fold2d.sum()) * len(job.t_axis_ms)

    run(1)                      # warm-up: library load, OpenMP thread start
    dt1, _ = run(ny)            # one inline: every thread busy
    n_x = int(max(1, min(tile.x1 - tile.x0 - ix0, round(budget_s / max(dt1, 1e-3)))))
    dt, c = run(ny, n_x)
    log(f"cpu baseline: {n_x} x {ny} pillars x {len(rows)} traces in {dt:.1f} s on {threads} threads")
    return {"value": round(c / dt / 1e9, 4), "unit": "Gcontrib/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "sample": (f"{n_x} x {ny} pillars (whole inlines of subset tile {tile_index}) x all {len(rows)} "
                       f"traces of that tile x {len(job.t_axis_ms)} samples, same settings and exact "
                       f"traveltime; C_sample={c}, {dt:.2f} s; oracle C++ FP64 -O3 -ffast-math OpenMP")}


def point_diffractor_plain(kernel, reps: int) -> dict:
    """The plain straight-ray gather on the 4096-trace point-diffractor job
    (64 x 64 x 500 output, one tile, aperture 1000 m, weights on, no AA):
    one plan replayed ``reps`` times, HIP events on the launch stream."""
    import torch

    from pm_amd import roofline
    from pm_amd.kernels import _abi
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import point_diffractor_job

    lib = _abi.load_library()
    job = point_diffractor_job(device="cuda")
    ex = TileExecutor(kernel, job)
    e = count_effective(ex, [0])
    plan = ex.plan()
    plan.run(use_graph=USE_GRAPH)  # capture
    torch.cuda.synchronize()
    ms = plan_launch_ms(lib, [plan], reps)
    ex.reset()
    plan.run(use_graph=USE_GRAPH)
    torch.cuda.synchronize()
    c = ex.contributions()
    info = plan.info()
    plan.destroy()
    f = roofline.flops_per_contribution("straight_ray", "none", True)
    out = roofline_entry(f, e, ms, "k_migrate (straight ray, weights, linear; one plan launch)")
    out.update({"workload": "point diffractor: 4096 traces x 1001 samples, 64x64 x 500 output, aperture 1000 m",
                "contributions": c, "gcontrib_s": round(c / (ms / 1e3) / 1e9, 2), "launch": info,
                "reps": reps})
    return out


def vti_headline(kernel, args) -> dict:
    """Config-4 job, full size; the learned-vs-exact comparison on the tile
    subset.  Returns the bench line's fields."""
    import torch
    from dataclasses import replace

    from pm_amd import roofline
    from pm_amd.kernels import _abi
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.lattice import Lattice
    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate, make_training_set
    from pm_amd.synthetic import vti_diffractor_job

    lib = _abi.load_library()
    t0 = time.perf_counter()
    job = vti_diffractor_job(n_traces=args.vti_traces, n_samples=2001, n_out=args.vti_out, nt=1500,
                             tile=args.region, device="cuda")
    torch.cuda.synchronize()
    log(f"vti job built in {time.perf_counter() - t0:.1f} s: {job.n_traces} traces x {job.n_samples}")
    tune = None
    if args.autotune:
        # output bins per tile (tile edge) x launch shape, on the exact
        # traveltime over the step's own region (the same work items per
        # workgroup as the timed steps: a larger region would favour shapes
        # whose item count does not divide evenly over the workgroups here)
        from pm_amd.autotune import autotune_tiles

        edges = tuple(e for e in (16, 32, 64) if e <= args.region)
        tune = autotune_tiles(kernel, job, edges=edges, region=args.region,
                              candidates=[(3, 16, 0), (4, 16, 0), (3, 16, 1)], max_edge=args.region)
        log(f"autotune: edge {tune.edge}, mw {tune.mw}, batch {tune.batch}, pipe {tune.pipe}")
    edge = tune.edge if tune else args.region
    job.tile = edge
    t0 = time.perf_counter()
    ex = TileExecutor(kernel, job)
    torch.cuda.synchronize()
    from pm_amd.autotune import region_tiles

    nx_, ny_, _ = ex.shape
    rx0, rx1, ry0, ry1 = central_region(nx_, ny_, args.region, args.region)
    sub = region_tiles(ex, rx0, rx1, ry0, ry1)
    log(f"executor ready in {time.perf_counter() - t0:.1f} s: {len(ex.tiles)} tiles of {edge}; step region "
        f"x{rx0}-{rx1 - 1} y{ry0}-{ry1 - 1}: tiles {sub} rows {[int(ex.tiles[i].rows.shape[0]) for i in sub]}")
    vc, ec = job.cube, job.meta["eta_cube"]
    d_max = job.config.max_aperture_m + 1500.0
    norm = Normalizer(tau_max=1.5, d_max=d_max, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
    t0 = time.perf_counter()
    tr = make_training_set(vc, ec, args.train_samples, d_max, norm, seed=1, t_max_ms=3000.0)
    torch.cuda.synchronize()
    set_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    sur = TraveltimeSurrogate.create(norm, seed=0)
    sur.weight_terms = args.weight_terms
    # an f32-trained network; the gather rounds its weights to bf16 (one
    # term) or splits them in two (--weight-terms 2)
    if USE_GRAPH:
        sur.train(tr, steps=args.train_steps, batch=args.train_batch, log_every=max(args.train_steps // 8, 1),
                  qat_fraction=args.qat, verbose=True)
    else:
        sur.train_stepwise(tr, steps=args.train_steps, batch=args.train_batch,
                           log_every=max(args.train_steps // 8, 1), qat_fraction=args.qat, verbose=True)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    held = sur.error_samples(tr, job.dt_ms, bf16=True) if args.train_samples <= 2_000_000 else None
    del tr
    log(f"surrogate: training set {set_s:.1f} s, training {train_s:.1f} s")
    K = Lattice(*(int(v) for v in args.lattice.split(",")))
    errs = ex.check_surrogate(sur, node_spacing=K, tiles=sub)
    if args.force_learned:
        for i in sub:
            ex.tiles[i].learned_ok = True
    learned = [i for i in sub if ex.tiles[i].learned_ok]
    fallback = [i for i in sub if not ex.tiles[i].learned_ok]
    log(f"held-out max error per subset tile (samples): {[round(e, 4) for e in errs]}; learned {learned}, "
        f"exact fallback {fallback}")

    # exact VTI traveltime on the subset: E, one captured run, one timed run
    e_exact = count_effective(ex, sub)
    exact_plan = ex.plan(tiles=sub)
    exact_plan.run(use_graph=USE_GRAPH)
    torch.cuda.synchronize()
    reset_tiles(ex, sub)
    exact_ms = plan_launch_ms(lib, [exact_plan])
    img_exact = subset_image(ex, sub)
    c_sub = subset_contributions(ex, sub)
    exact_info = exact_plan.info()
    exact_plan.destroy()
    log(f"exact VTI: {exact_ms:.1f} ms, C = {c_sub:.4e}, E = {e_exact:.4e}")

    # the method without the network: isotropic straight-ray time in the same
    # v_rms (also the plain straight-ray gather's measurement on this subset)
    saved = ex.config
    ex.config = replace(saved, kernel_type="straight_ray", vti_enabled=False)
    e_plain = count_effective(ex, sub)
    iso_plan = ex.plan(tiles=sub)
    iso_plan.run(use_graph=USE_GRAPH)
    torch.cuda.synchronize()
    reset_tiles(ex, sub)
    plain_ms = plan_launch_ms(lib, [iso_plan])
    img_iso = subset_image(ex, sub)
    plain_info = iso_plan.info()
    iso_plan.destroy()
    ex.config = saved
    reset_tiles(ex, sub)

    # learned traveltime: the timed steps
    e_learned = count_effective(ex, learned, traveltime=_abi.TT_VTI_LEARNED, surrogate=sur,
                                node_spacing=K) if learned else 0
    plans = []
    if learned:
        plans.append(ex.plan(tiles=learned, traveltime=_abi.TT_VTI_LEARNED, surrogate=sur, node_spacing=K))
    if fallback:
        plans.append(ex.plan(tiles=fallback))
    use_graph = USE_GRAPH
    for _ in range(args.warmup):
        reset_tiles(ex, sub)
        for p in plans:
            p.run(use_graph=use_graph)   # the first run captures the graph
    torch.cuda.synchronize()
    lib.pstm_hip_profile_reset(1)
    t0 = time.perf_counter()
    for s in range(args.steps):
        reset_tiles(ex, sub)
        for p in plans:
            p.run(use_graph=use_graph)
        torch.cuda.synchronize()
        log(f"step {s} done")
    elapsed = time.perf_counter() - t0
    kms, nl = ctypes.c_double(), ctypes.c_longlong()
    lib.pstm_hip_profile_read(ctypes.byref(kms), ctypes.byref(nl))
    lib.pstm_hip_profile_reset(0)
    learned_step_ms = kms.value / max(args.steps, 1)
    img_learned = subset_image(ex, sub)
    learned_info = plans[0].info()
    for p in plans:
        p.destroy()
    ms_step = elapsed * 1000.0 / max(args.steps, 1)
    nrm = float(torch.linalg.norm(img_exact))
    rel = float(torch.linalg.norm(img_learned - img_exact)) / nrm
    rel_iso = float(torch.linalg.norm(img_iso - img_exact)) / nrm

    # roofline of the learned gather (the learned plan's launches alone when
    # some tiles fell back: timed separately)
    f_learned = roofline.FLOPS_PER_CONTRIBUTION["vti_learned"]
    if learned and not fallback:
        rl = roofline_entry(f_learned, e_learned, learned_step_ms,
                            "k_migrate<VTI_LEARNED> (fused BF16-MFMA lattice inference + VALU gather)")
    elif learned:
        lp = ex.plan(tiles=learned, traveltime=_abi.TT_VTI_LEARNED, surrogate=sur, node_spacing=K)
        lp.run(use_graph=USE_GRAPH)
        torch.cuda.synchronize()
        rl = roofline_entry(f_learned, e_learned, plan_launch_ms(lib, [lp]),
                            "k_migrate<VTI_LEARNED> (learned tiles only)")
        lp.destroy()
    else:
        rl = roofline_entry(roofline.FLOPS_PER_CONTRIBUTION["vti"], e_exact, learned_step_ms,
                            "k_migrate<VTI> (every subset tile fell back to the exact traveltime)")
    rl["launches_timed"] = int(nl.value)
    reset_tiles(ex, sub)
    tiles_desc = f"x{rx0}-{rx1 - 1} y{ry0}-{ry1 - 1} in {len(sub)} tile(s) of {edge}x{edge}"
    res = {
        "value": c_sub / (elapsed / max(args.steps, 1)) / 1e9, "ms_per_step": ms_step,
        "config": {
            "workload": (f"VTI diffractors (learned vs exact traveltime): {job.n_traces} traces x 2001 samples "
                         f"(2 ms), 300 diffractors, 3-D v_rms 1500-4500 m/s and eta 0-0.2 on 128x128x201 grids, "
                         f"output {args.vti_out}x{args.vti_out} bins at 12.5 m x 1500 samples, aperture 2500 m, "
                         f"dip 60, weights on; the whole job resident in HBM, a step migrates the central "
                         f"{args.region}x{args.region} pillars ({tiles_desc})"),
            "traces": job.n_traces, "samples": job.n_samples, "output": list(ex.shape),
            "tiles_total": len(ex.tiles), "tiles_per_step": len(sub), "tile": edge,
            "pillars_per_step": args.region * args.region,
            "contributions_per_step": c_sub, "effective_contributions_per_step_learned": e_learned,
            "traveltime": "learned VTI (MLP 5-64x4-2 SiLU, BF16 MFMA, FP32 accumulation) with per-tile exact "
                          "fallback", "parallelism": "1 GPU",
            "launch": "one persistent hipGraph-captured gather per plan per step"},
        "roofline": rl,
        "learned_vs_exact": {
            "exact_ms": round(exact_ms, 2), "learned_ms": round(learned_step_ms, 2),
            "learned_speedup": round(exact_ms / learned_step_ms, 3),
            "exact_gcontrib_s": round(c_sub / (exact_ms / 1e3) / 1e9, 2),
            "learned_gcontrib_s": round(c_sub / (learned_step_ms / 1e3) / 1e9, 2),
            "image_rel_l2_learned_vs_exact": rel, "image_rel_l2_without_network_vs_exact": rel_iso,
            "effective_contributions": {"exact": e_exact, "learned": e_learned},
            "exact_roofline": roofline_entry(roofline.FLOPS_PER_CONTRIBUTION["vti"], e_exact, exact_ms,
                                             "k_migrate<VTI> (exact VTI traveltime)"),
            "launch_shape": {"learned": learned_info, "exact": exact_info},
            "surrogate": {"train_samples": args.train_samples, "steps": args.train_steps,
                          "batch": args.train_batch, "train_s": round(train_s, 2),
                          "training_set_s": round(set_s, 2),
                          "lattice": {"spacing": K.spacing, "growth": K.growth, "max_spacing": K.max_spacing,
                                      "nodes_per_trace": int(len(K.nodes(ex.nt)))},
                          "heldout_max_samples_per_tile": [round(e, 4) for e in errs],
                          "tiles_learned": len(learned), "tiles_exact_fallback": len(fallback),
                          "training_set_error_samples": held, "forced": bool(args.force_learned),
                          "weight_terms": args.weight_terms, "qat_fraction": args.qat},
        },
        "plain_straight_ray": dict(roofline_entry(roofline.FLOPS_PER_CONTRIBUTION["straight"], e_plain, plain_ms,
                                                  "k_migrate (straight ray, weights, linear) on the same subset"),
                                   gcontrib_s=round(c_sub / (plain_ms / 1e3) / 1e9, 2), launch=plain_info),
    }
    if tune is not None:
        res["autotune"] = tune.as_dict()
    res["_ex"] = ex
    res["_sub"] = sub
    return res


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


def dropin_section(kernel, ex, lib) -> dict:
    """The reference's call path on the busiest tile: HipKernel.migrate_tile
    with host NumPy arrays (TraceBlock, OutputTile, 3-D VelocitySlice), which
    uploads the tile's traces, migrates and downloads the tile, timed by the
    host clock; beside it the upload alone and the device-resident plan over
    the same tile (HIP events)."""
    import numpy as np
    import torch

    from pm_amd.kernels.base import OutputTile, VelocitySlice, create_trace_block
    from pm_amd.kernels.hip_backend import DeviceTraces

    i = max(range(len(ex.tiles)), key=lambda k: int(ex.tiles[k].rows.shape[0]))
    t = ex.tiles[i]
    h = ex.host_traces(t.rows.cpu().numpy().astype(np.int64))
    job = ex.job
    tb = create_trace_block(h["amplitudes"], h["source_x"], h["source_y"], h["receiver_x"], h["receiver_y"],
                            sample_rate_ms=job.dt_ms)
    xa, ya = t.x_axis.cpu().numpy(), t.y_axis.cpu().numpy()
    nt = len(job.t_axis_ms)
    vel = VelocitySlice(vrms=t.vrms.cpu().numpy().astype(np.float64), is_1d=False)

    def tile():
        return OutputTile(image=np.zeros((len(xa), len(ya), nt)), fold=np.zeros((len(xa), len(ya)), np.int32),
                          x_axis=xa, y_axis=ya, t_axis_ms=np.asarray(job.t_axis_ms, float))

    kernel.migrate_tile(tb, tile(), vel, ex.config)   # warm-up
    out = tile()
    t0 = time.perf_counter()
    kernel.migrate_tile(tb, out, vel, ex.config)
    call_s = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dt = DeviceTraces.from_block(tb, ex.device)
    torch.cuda.synchronize()
    up_s = time.perf_counter() - t0
    del dt
    plan = ex.plan(tiles=[i])
    ex.reset()
    plan.run(use_graph=USE_GRAPH)
    torch.cuda.synchronize()
    plan_ms = plan_launch_ms(lib, [plan])
    plan.destroy()
    ex.reset()
    c = int(out.fold.sum()) * nt
    gb = tb.amplitudes.nbytes / 1e9
    log(f"drop-in: tile {i} ({tb.n_traces} traces, {gb:.2f} GB) {call_s:.3f} s per call, upload {up_s:.3f} s, "
        f"plan {plan_ms:.0f} ms")
    return {"tile": i, "traces": int(tb.n_traces), "trace_gb": round(gb, 3), "call_s": round(call_s, 4),
            "upload_s": round(up_s, 4), "upload_gbps": round(gb / up_s, 1), "device_plan_ms": round(plan_ms, 2),
            "contributions": c, "gcontrib_s_pcie_inclusive": round(c / call_s / 1e9, 2),
            "gcontrib_s_device": round(c / (plan_ms / 1e3) / 1e9, 2)}


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
    tune = autotune(kernel, ex, candidates=[(2, 16, 0), (2, 8, 0), (1, 16, 0), (2, 16, 1)])
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


def reflector_section(kernel, args, sections) -> dict:
    """The anti-aliased reflector job (200,000 traces x 1501 samples, 3-D
    v_rms, 256 x 256 x 1000 output in 16 tiles, triangle anti-aliasing):
    one pass over every tile as one plan, timed with HIP events, and the
    side sections that run on it."""
    import torch

    from pm_amd import roofline
    from pm_amd.kernels import _abi
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import reflector_job

    lib = _abi.load_library()
    lib.pstm_hip_set_tuning(0, 0)
    lib.pstm_hip_set_pipeline(-1)
    job = reflector_job(device="cuda")
    ex = TileExecutor(kernel, job)
    all_tiles = list(range(len(ex.tiles)))
    e = count_effective(ex, all_tiles)
    plan = ex.plan()
    ex.reset()
    plan.run(use_graph=USE_GRAPH)
    torch.cuda.synchronize()
    ex.reset()
    ms = plan_launch_ms(lib, [plan])
    c = ex.contributions()
    out = roofline_entry(roofline.FLOPS_PER_CONTRIBUTION["triangle_aa"], e, ms,
                         "k_migrate (straight ray + triangle AA, one plan over 16 tiles)")
    out.update({"workload": ("reflectors: 200000 traces x 1501 samples, 40 planes + 200 diffractors, 3-D v_rms "
                             "64x64x151, output 256x256 x 1000 in 16 tiles of 64x64, aperture 2500 m, dip 60, "
                             "triangle AA"), "contributions": c, "gcontrib_s": round(c / (ms / 1e3) / 1e9, 2),
                "launch": plan.info()})
    plan.destroy()
    log(f"reflectors: {ms:.0f} ms, {out['gcontrib_s']} Gcontrib/s")
    if "output" in sections:
        out["output"] = output_section(kernel, ex)
    if "time_variant" in sections:
        img_u, _ = ex.assemble()
        out["time_variant"] = time_variant_section(kernel, job, img_u, ms, e)
        del img_u
    if "interpolation" in sections:
        out["interpolation"] = interpolation_section(ex)
    if "dropin" in sections:
        out["dropin"] = dropin_section(kernel, ex, lib)
    lib.pstm_hip_set_tuning(0, 0)
    lib.pstm_hip_set_pipeline(-1)
    del ex
    torch.cuda.empty_cache()
    return out


SECTIONS = ("reflectors", "output", "time_variant", "interpolation", "dropin", "stream")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--region", type=int, default=32,
                    help="pillars per side of the central output region a step migrates")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the tile-edge x launch-shape autotuner (built-in shape, tile edge = region)")
    ap.add_argument("--vti-traces", type=int, default=500_000)
    ap.add_argument("--vti-out", type=int, default=512)
    ap.add_argument("--train-samples", type=int, default=10_000_000)
    ap.add_argument("--train-steps", type=int, default=100_000)
    ap.add_argument("--train-batch", type=int, default=65_536)
    ap.add_argument("--weight-terms", type=int, default=1, choices=[1, 2],
                    help="fused inference weights of the f32-trained network: 1 rounded to bf16, 2 as two bf16 "
                         "terms (one more MFMA per product)")
    ap.add_argument("--qat", type=float, default=0.25,
                    help="fraction of the training steps (the last ones) run through the fused inference's bf16 "
                         "numerics (quantisation-aware)")
    ap.add_argument("--lattice", default="8,100,64",
                    help="t0 lattice of the learned traveltime: smallest spacing, growth (1/256), largest spacing")
    ap.add_argument("--force-learned", action="store_true",
                    help="use the learned traveltime on every subset tile whatever its held-out error "
                         "(profiling runs with a briefly trained network)")
    ap.add_argument("--sections-only", action="store_true",
                    help="profiler runs of the side sections: skip the VTI headline, run --sections only")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: the learned plans launched directly (no graph), no CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample (0 = skip)")
    ap.add_argument("--plain-reps", type=int, default=20,
                    help="replays of the point-diffractor plan for the plain-kernel roofline (0 = skip)")
    ap.add_argument("--sections", default="",
                    help="comma list of side measurements on the reflector job: " + ", ".join(SECTIONS))
    ap.add_argument("--stream-bins", type=int, default=60)
    ap.add_argument("--stream-traces", type=int, default=240_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the migration shards "
              "by output tile with no collective, so N GPUs run N independent tile sets",
              file=sys.stderr)
        sys.exit(2)
    sections = {s for s in args.sections.split(",") if s}
    unknown = sections - set(SECTIONS)
    if unknown:
        raise SystemExit(f"unknown sections {sorted(unknown)}; choose from {SECTIONS}")

    import torch

    from pm_amd.kernels import HipKernel
    from pm_amd.kernels import KernelConfig
    from pm_amd.kernels import _abi

    global USE_GRAPH
    USE_GRAPH = not args.profile
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    kernel = HipKernel()
    kernel.initialize(KernelConfig())
    t_all = time.perf_counter()
    if args.profile:
        calibration_kernel()
    if args.sections_only:
        out = {"sections_only": True}
        if sections & {"reflectors", "output", "time_variant", "interpolation", "dropin"}:
            out["reflectors"] = reflector_section(kernel, args, sections)
        if "stream" in sections:
            out["streamed"] = streamed_section(kernel, args)
        out["wall_s"] = round(time.perf_counter() - t_all, 1)
        print(json.dumps(out), flush=True)
        kernel.cleanup()
        return
    head = vti_headline(kernel, args)
    ex, sub = head.pop("_ex"), head.pop("_sub")
    if args.profile:
        log("profile run done")
        kernel.cleanup()
        return
    cpu = None
    if args.cpu_seconds > 0:
        busiest = max(sub, key=lambda i: int(ex.tiles[i].rows.shape[0]))
        cpu = cpu_baseline(ex, busiest, args.cpu_seconds, _abi.TT_VTI)
    del ex
    torch.cuda.empty_cache()
    plain_pd = point_diffractor_plain(kernel, args.plain_reps) if args.plain_reps > 0 else None
    out = {
        "metric": METRIC, "value": round(head["value"], 3), "unit": "Gcontrib/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms_per_step"], 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": head["config"], "roofline": head["roofline"], "cpu_baseline": cpu,
        "learned_vs_exact": head["learned_vs_exact"],
        "plain_straight_ray": {"vti_subset": head["plain_straight_ray"], "point_diffractor": plain_pd},
    }
    if "autotune" in head:
        out["autotune"] = head["autotune"]
    if cpu:
        out["gpu_over_cpu"] = {"learned": round(head["value"] / cpu["value"], 1),
                               "exact": round(head["learned_vs_exact"]["exact_gcontrib_s"] / cpu["value"], 1)}
    if sections & {"reflectors", "output", "time_variant", "interpolation", "dropin"}:
        out["reflectors"] = reflector_section(kernel, args, sections)
    if "stream" in sections:
        out["streamed"] = streamed_section(kernel, args)
    out["wall_s"] = round(time.perf_counter() - t_all, 1)
    print(json.dumps(out), flush=True)
    kernel.cleanup()


if __name__ == "__main__":
    main()
