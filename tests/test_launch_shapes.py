"""Launch shapes of the gather: sample blocks per wave (MW), traces per LDS
batch, and pipelined staging.  Pipelined staging keeps the next batch's f32
rows in registers; for the f64 anti-aliasing rows it loads them straight from
global memory into a second LDS buffer.  Unpipelined staging (the default)
loads the current batch's rows, f32 or f64, the same direct way into the one
buffer at the batch start.  A shape changes only
how the f32 per-batch sums are grouped, so every shape must give the FP64
oracle's image within the FP32 bar and the oracle's fold exactly."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from pm_amd.kernels import _abi

pytestmark = pytest.mark.gpu

# (MW cap, traces-per-batch cap, pipelined staging)
SHAPES = [(4, 7, 0), (4, 3, 1), (2, 10, 0), (1, 5, 1), (3, 16, 0)]


@pytest.fixture(scope="module")
def kernel():
    from pm_amd.kernels import HipKernel, KernelConfig

    k = HipKernel()
    k.initialize(KernelConfig())
    yield k
    k.cleanup()


@pytest.mark.parametrize("antialias", ["none", "triangle"])
def test_every_launch_shape_matches_oracle(kernel, antialias):
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import reflector_job

    # four 10 x 10 tiles whose image holds focused events, not only migrated
    # noise (on noise alone the f32 error sits at the bar)
    job = reflector_job(seed=8, n_traces=5000, n_samples=501, n_out=20, nt=400, d_out=25.0, tile=10,
                        n_planes=6, n_diffractors=10, antialias=antialias, device="cuda")
    ex = TileExecutor(kernel, job)
    t = max(ex.tiles, key=lambda t: int(t.rows.shape[0]))
    h = ex.host_traces(t.rows.cpu().numpy().astype(np.int64))
    c = job.config
    ref = orc.migrate(orc.MigrateInputs(
        amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
        receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
        midpoint_y=h["midpoint_y"], x_axis=t.x_axis.cpu().numpy(), y_axis=t.y_axis.cpu().numpy(),
        t_axis_ms=job.t_axis_ms, vrms=t.vrms.cpu().numpy(), dt_in_ms=job.dt_ms,
        max_aperture_m=c.max_aperture_m, min_aperture_m=c.min_aperture_m, max_dip_deg=c.max_dip_degrees,
        taper_fraction=c.taper_fraction, apply_spreading=c.apply_spreading,
        apply_obliquity=c.apply_obliquity, grid_dx=c.grid_dx, grid_dy=c.grid_dy,
        aa_mode=_abi.AA_TRIANGLE if antialias == "triangle" else _abi.AA_NONE,
        aa_max_half_width=c.aa_max_half_width))
    assert np.linalg.norm(ref.image) > 0
    lib = _abi.load_library()
    try:
        for mw, batch, pipe in SHAPES:
            assert lib.pstm_hip_set_tuning(mw, batch) == 0 and lib.pstm_hip_set_pipeline(pipe) == 0
            plan = ex.plan()
            info = plan.info()
            ex.reset()
            plan.run(use_graph=False)
            img, fold = ex.assemble()
            plan.destroy()
            got = img[t.x0:t.x1, t.y0:t.y1].double().cpu().numpy()
            rel = np.linalg.norm(got - ref.image) / np.linalg.norm(ref.image)
            print(f"{antialias} mw<={mw} batch<={batch} pipe={pipe} -> {info}: rel-L2 {rel:.2e}")
            assert rel <= 1e-5, (mw, batch, pipe)
            np.testing.assert_array_equal(fold[t.x0:t.x1, t.y0:t.y1].cpu().numpy(), ref.fold2d)
    finally:
        lib.pstm_hip_set_tuning(0, 0)
        lib.pstm_hip_set_pipeline(-1)
        ex.reset()
