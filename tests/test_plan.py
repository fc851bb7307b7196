"""Plans: many tiles (and common-offset bins) in one persistent,
hipGraph-captured gather launch fed by a device work queue.

* A plan run equals the tile-by-tile runs bit for bit (same per-item
  arithmetic and order), eagerly and from the captured graph, over several
  replays with eager work between them.
* E counted through a plan equals the tile-by-tile count.
* Common-offset bins with uneven counts migrate into one image in one launch,
  within the FP32 bar of the FP64 oracle on a tile, fold bit-exact.
* The learned traveltime runs through a plan as well.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from pm_amd.kernels import _abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kernel():
    from pm_amd.kernels import HipKernel, KernelConfig

    k = HipKernel()
    k.initialize(KernelConfig())
    yield k
    k.cleanup()


def _state(ex):
    import torch

    img, fold = ex.assemble()
    torch.cuda.synchronize()
    return img.clone(), fold.clone()


@pytest.mark.parametrize("antialias", ["none", "triangle"])
def test_plan_equals_tile_runs(kernel, antialias):
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import reflector_job

    job = reflector_job(n_traces=12_000, n_samples=1001, n_out=48, nt=400, tile=16, antialias=antialias,
                        n_planes=10, n_diffractors=20, device="cuda")
    ex = TileExecutor(kernel, job)
    ex.reset()
    ex.run()
    ref_img, ref_fold = _state(ex)
    assert float(ref_img.abs().max()) > 0
    plan = ex.plan()
    info = plan.info()
    print("plan", info, "tiles", len(ex.tiles))
    assert info["items"] > 0 and info["grid"] > 0
    ex.reset()
    plan.run(use_graph=False)
    img, fold = _state(ex)
    assert torch.equal(img, ref_img) and torch.equal(fold, ref_fold)
    for rep in range(3):
        ex.reset()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        plan.run(use_graph=True)
        img, fold = _state(ex)
        assert bool(torch.isfinite(junk))
        assert torch.equal(img, ref_img), f"replay {rep}"
        assert torch.equal(fold, ref_fold), f"replay {rep}"
    # accumulation: two runs without a reset double the image and fold
    ex.reset()
    plan.run()
    plan.run()
    img, fold = _state(ex)
    assert torch.equal(fold, 2 * ref_fold)
    assert torch.allclose(img, 2 * ref_img, rtol=1e-6, atol=1e-6 * float(ref_img.abs().max()))
    plan.destroy()


def test_plan_counts_effective(kernel):
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import reflector_job

    job = reflector_job(n_traces=6000, n_samples=801, n_out=32, nt=300, tile=16, antialias="none",
                        n_planes=5, n_diffractors=10, device="cuda")
    ex = TileExecutor(kernel, job)
    ex.reset()
    e_tiles = ex.run(count_effective=True)
    plan = ex.plan(count_effective=True)
    ex.reset()
    plan.run()
    kernel.synchronize()
    assert int(plan.counter.item()) == e_tiles > 0
    plan.run()
    kernel.synchronize()
    assert int(plan.counter.item()) == 2 * e_tiles
    plan.destroy()


def test_bins_one_launch_against_oracle(kernel):
    """Common-offset bins with lognormal counts, one plan over every tile; one
    tile checked against the FP64 oracle over all bins' traces."""
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import bins_job

    job = bins_job(total=8000, n_samples=801, n_out=32, nt=300, tile=16, n_planes=8, n_diffractors=10,
                   device="cuda")
    counts = np.diff(job.meta["bin_start"])
    assert len(counts) == 40 and counts.max() > 3 * counts.min()
    ex = TileExecutor(kernel, job)
    plan = ex.plan()
    ex.reset()
    plan.run()
    img, fold = _state(ex)
    t = max(ex.tiles, key=lambda t: int(t.rows.shape[0]))
    rows = t.rows.cpu().numpy().astype(np.int64)
    h = ex.host_traces(rows)
    cfg = ex.config
    inp = orc.MigrateInputs(
        amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
        receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
        midpoint_y=h["midpoint_y"], x_axis=t.x_axis.cpu().numpy(), y_axis=t.y_axis.cpu().numpy(),
        t_axis_ms=job.t_axis_ms, vrms=t.vrms.cpu().numpy(), dt_in_ms=job.dt_ms,
        max_aperture_m=cfg.max_aperture_m, min_aperture_m=cfg.min_aperture_m, max_dip_deg=cfg.max_dip_degrees,
        taper_fraction=cfg.taper_fraction, apply_spreading=cfg.apply_spreading,
        apply_obliquity=cfg.apply_obliquity, grid_dx=cfg.grid_dx, grid_dy=cfg.grid_dy)
    ref = orc.migrate(inp)
    got = img[t.x0:t.x1, t.y0:t.y1].double().cpu().numpy()
    rel = np.linalg.norm(got - ref.image) / np.linalg.norm(ref.image)
    print("bins tile rel-L2 vs FP64 oracle", rel)
    assert rel <= 1e-5
    np.testing.assert_array_equal(fold[t.x0:t.x1, t.y0:t.y1].cpu().numpy(), ref.fold2d)
    plan.destroy()


def test_learned_plan_equals_tile_runs(kernel):
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate
    from pm_amd.synthetic import vti_diffractor_job

    job = vti_diffractor_job(n_traces=6000, n_samples=801, n_out=32, nt=300, tile=16, n_diffractors=10,
                             device="cuda")
    ex = TileExecutor(kernel, job)
    sur = TraveltimeSurrogate.create(Normalizer(0.3, 4000.0, 1500.0, 4500.0, 0.2), seed=1)
    for t in ex.tiles:
        t.learned_ok = True
    ex.reset()
    ex.run(surrogate=sur, node_spacing=16)
    ref_img, ref_fold = _state(ex)
    plan = ex.plan(traveltime=_abi.TT_VTI_LEARNED, surrogate=sur, node_spacing=16)
    ex.reset()
    plan.run()
    img, fold = _state(ex)
    assert torch.equal(img, ref_img) and torch.equal(fold, ref_fold)
    plan.destroy()


@pytest.mark.parametrize("kind", ["vti", "curved"])
def test_plan_other_traveltimes_equal_tile_runs(kernel, kind):
    """Exact VTI (3-D eta slabs per tile) and curved-ray plans equal their
    tile-by-tile launches bit for bit."""
    import torch
    from dataclasses import replace

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import reflector_job, vti_diffractor_job

    if kind == "vti":
        job = vti_diffractor_job(n_traces=6000, n_samples=801, n_out=32, nt=300, tile=16, n_diffractors=10,
                                 device="cuda")
    else:
        job = reflector_job(n_traces=6000, n_samples=801, n_out=32, nt=300, tile=16, antialias="none",
                            n_planes=5, n_diffractors=10, device="cuda")
        job.config = replace(job.config, kernel_type="curved_ray", curved_ray_enabled=True, curved_ray_v0=1600.0,
                             curved_ray_k=0.4)
    ex = TileExecutor(kernel, job)
    assert kind != "vti" or all(t.eta is not None for t in ex.tiles)
    ex.reset()
    ex.run()
    ref_img, ref_fold = _state(ex)
    assert float(ref_img.abs().max()) > 0
    plan = ex.plan()
    ex.reset()
    plan.run()
    img, fold = _state(ex)
    assert torch.equal(img, ref_img) and torch.equal(fold, ref_fold)
    plan.destroy()


def test_temporary_plan_equals_kept_plan(kernel):
    """``ex.plan().run()`` frees the plan as soon as run() returns, while its
    graph may still be executing: destruction waits for it, so the image
    equals that of a plan kept until after synchronisation, also when other
    device allocations follow at once."""
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import reflector_job

    job = reflector_job(seed=9, n_traces=6000, n_samples=601, n_out=24, nt=500, d_out=25.0, tile=12,
                        n_planes=6, n_diffractors=10, antialias="none", device="cuda")
    ex = TileExecutor(kernel, job)
    plan = ex.plan()
    plan.run()
    torch.cuda.synchronize()
    ref, _ = _state(ex)
    plan.destroy()
    ex.reset()
    ex.plan().run()                       # temporary: destroyed on return
    scratch = [torch.full((1 << 24,), float(i), device="cuda") for i in range(4)]
    img, _ = _state(ex)
    del scratch
    assert torch.equal(img, ref)
