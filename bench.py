"""Benchmark: Gcontrib/s of the MI355X PSTM backend on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload reflectors]

A step is one migration pass over the whole synthetic job with every input
already resident in HBM: for the default workload, 16 tile launches (64 x 64
pillars x 1000 samples each) over 200,000 traces of 1501 samples with 3-D
v_rms and triangle anti-aliasing.  A contribution is one (trace, output
sample) pair inside the maximum aperture, C = sum of the 3-D fold; the rate
is C / step time.  The roofline entry uses E (contributions that pass the
time-dependent aperture and the sample bounds) x the frozen FLOPs per
contribution over the gather kernel's own launch time, measured with HIP
events on the launch stream.  The CPU baseline is the C++ oracle (built with
-O3 -ffast-math, OpenMP) on a bounded pillar sample of the same job.

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

    def run(n_p: int):
        inp = orc.MigrateInputs(
            amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
            receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
            midpoint_y=h["midpoint_y"], x_axis=xa[ix:ix + 1], y_axis=ya[:n_p],
            t_axis_ms=job.t_axis_ms, vrms=(v3[ix:ix + 1, :n_p] if v3 is not None else v1),
            dt_in_ms=job.dt_ms, max_aperture_m=cfg.max_aperture_m,
            min_aperture_m=cfg.min_aperture_m, max_dip_deg=cfg.max_dip_degrees,
            taper_fraction=cfg.taper_fraction, apply_spreading=cfg.apply_spreading,
            apply_obliquity=cfg.apply_obliquity, grid_dx=cfg.grid_dx, grid_dy=cfg.grid_dy,
            aa_mode=_abi.AA_TRIANGLE if cfg.antialias == "triangle" else _abi.AA_NONE,
            aa_max_half_width=cfg.aa_max_half_width)
        t0 = time.perf_counter()
        res = orc.migrate(inp, threads=threads, fast=True)
        return time.perf_counter() - t0, int(res.fold2d.sum()) * len(job.t_axis_ms), res

    dt1, _, _ = run(1)
    ny = tile.y1 - tile.y0
    n_p = int(max(1, min(ny, round(budget_s / max(dt1, 1e-3)))))
    dt, c, _ = run(n_p)
    log(f"cpu baseline: {n_p} pillars x {len(rows)} traces in {dt:.1f} s on {threads} threads")
    return {"value": c / dt / 1e9, "unit": "Gcontrib/s", "cores": threads, "kind": "port",
            "sample": (f"{n_p} pillars (one inline of the busiest tile) x all {len(rows)} "
                       f"traces of that tile x {len(job.t_axis_ms)} samples, same settings; "
                       f"C_sample={c}, {dt:.2f} s; oracle C++ FP64 -O3 -ffast-math OpenMP")}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="reflectors", choices=["reflectors", "point_diffractor"])
    ap.add_argument("--autotune", action="store_true",
                    help="time candidate launch shapes on the busiest tile first")
    ap.add_argument("--candidates", default="",
                    help="autotune candidates as MWxBATCH,MWxBATCH,... (default: a 4x3 grid)")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: one untimed copy of 256 MiB (traffic calibration) "
                         "and the timed steps only, no counting pass")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample (0 = skip)")
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
    if args.autotune:
        from pm_amd.autotune import autotune

        cands = None
        if args.candidates:
            cands = [tuple(int(v) for v in c.split("x")) for c in args.candidates.split(",")]
        tune = autotune(kernel, ex, candidates=cands)
        log(f"autotune: mw={tune.mw} batch={tune.batch} ({tune.ms:.1f} ms on the busiest tile)")

    if args.profile:
        # plain copy of known size: reference point for the HBM byte counters
        a = torch.empty(64 * 2**20, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 268435456 bytes read, 268435456 bytes written")
        del a, b
        e_total = 0
        c_total = 0
    else:
        # E (counting pass, not timed) and C
        ex.reset()
        e_total = ex.run(count_effective=True)
        c_total = ex.contributions()
        img, _ = ex.assemble()
        assert bool(torch.isfinite(img).all()) and float(img.abs().max()) > 0, "empty or non-finite image"
        log(f"C = {c_total:.4e} contributions, E = {e_total:.4e}")

    for _ in range(args.warmup):
        ex.reset()
        ex.run()
    torch.cuda.synchronize()
    lib = _abi.load_library()
    lib.pstm_hip_profile_reset(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ex.reset()
        ex.run()
        log("step done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kms = ctypes.c_double()
    nl = ctypes.c_longlong()
    lib.pstm_hip_profile_read(ctypes.byref(kms), ctypes.byref(nl))
    lib.pstm_hip_profile_reset(0)
    ms_step = elapsed * 1000.0 / args.steps
    value = c_total / (elapsed / args.steps) / 1e9
    if args.profile:
        log(f"profile run: {ms_step:.1f} ms per step")
        kernel.cleanup()
        return

    cfg = job.config
    f = roofline.flops_per_contribution(cfg.kernel_type, cfg.antialias,
                                        bool(cfg.apply_spreading or cfg.apply_obliquity))
    launches = int(nl.value)
    launch_ms = kms.value / max(launches, 1)
    e_per_launch = e_total / len(ex.tiles)
    achieved = f * e_per_launch / (launch_ms / 1000.0) / 1e12
    peak = roofline.MI355X_FP32_VALU_TFLOPS
    rl = {"bound": "valu", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
          "frac": round(achieved / peak, 4), "traffic": None,
          "kernel": "k_migrate (gather)", "flops_per_contribution": f,
          "effective_contributions_per_launch": int(e_per_launch),
          "avg_launch_ms": round(launch_ms, 3), "launches_timed": launches,
          "frac_of_unpacked_fp32": round(achieved / roofline.MI355X_FP32_VALU_UNPACKED_TFLOPS, 4)}
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
                   "traveltime": "exact straight-ray DSR", "parallelism": "1 GPU"},
        "roofline": rl, "cpu_baseline": cpu,
        "launch": _abi.last_launch(lib),
    }
    if tune is not None:
        out["autotune"] = tune.as_dict()
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    kernel.cleanup()


if __name__ == "__main__":
    main()
