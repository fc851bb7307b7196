"""Trace interpolation methods (KernelConfig.interpolation_method, honoured by
the reference's NumbaKernel through interpolate_sample,
pstm/kernels/interpolation.py:276-318).

CPU: the oracle's interpolators against golden vectors written from the
reference's own functions (tests/golden/make_golden.py) and hand-derived
values; the method-name mapping.  GPU: migration with every method against the
FP64 oracle, and the combinations the backend refuses."""

from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from oracle import oracle as orc
from pm_amd.kernels import _abi

# reference code (get_method_code) -> code of this ABI
REF_TO_ABI = {0: _abi.INTERP_NEAREST, 1: _abi.INTERP_LINEAR, 2: _abi.INTERP_CUBIC, 3: _abi.INTERP_SINC4,
              4: _abi.INTERP_SINC8, 5: _abi.INTERP_SINC16, 6: _abi.INTERP_LANCZOS3, 7: _abi.INTERP_LANCZOS5}
METHODS = ["nearest", "linear", "cubic", "sinc4", "sinc8", "sinc16", "lanczos3", "lanczos5"]


def test_oracle_interpolators_match_golden(golden_dir):
    g = np.load(golden_dir / "interp_methods.npz")
    tr, pos, val = g["trace"], g["pos"], g["value"]
    for rc in g["reference_code"]:
        got = np.array([orc.interp_sample(tr, t, REF_TO_ABI[int(rc)]) for t in pos])
        np.testing.assert_allclose(got, val[rc], rtol=0, atol=1e-13, err_msg=f"method {rc}")


def test_oracle_interpolators_by_hand():
    tr = np.array([0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0], np.float32)
    # Catmull-Rom, textbook form 0.5 (2 p1 + (p2 - p0) t + (2 p0 - 5 p1 + 4 p2 - p3) t^2
    # + (3 p1 - p0 - 3 p2 + p3) t^3) with p = (0, 0, 1, 0) at t = 0.25 (position 1.25)
    p0, p1, p2, p3, t = 0.0, 0.0, 1.0, 0.0, 0.25
    cr = 0.5 * (2 * p1 + (p2 - p0) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t ** 2
                + (3 * p1 - p0 - 3 * p2 + p3) * t ** 3)
    assert orc.interp_sample(tr, 1.25, _abi.INTERP_CUBIC) == pytest.approx(cr, abs=1e-15)
    # nearest rounds half up
    assert orc.interp_sample(tr, 1.5, _abi.INTERP_NEAREST) == 1.0
    assert orc.interp_sample(tr, 1.49, _abi.INTERP_NEAREST) == 0.0
    ramp = np.arange(30, dtype=np.float32)
    const = np.full(30, 2.5, np.float32)
    for code in (_abi.INTERP_SINC4, _abi.INTERP_SINC8, _abi.INTERP_SINC16, _abi.INTERP_LANCZOS3,
                 _abi.INTERP_LANCZOS5, _abi.INTERP_CUBIC):
        # a node returns its sample; a constant trace stays constant (weights normalised)
        assert orc.interp_sample(ramp, 12.0, code) == pytest.approx(12.0, abs=1e-12)
        assert orc.interp_sample(const, 13.37, code) == pytest.approx(2.5, abs=1e-12)
        # near the ends every wide kernel falls back to linear interpolation
        assert orc.interp_sample(ramp, 0.3, code) == pytest.approx(0.3, abs=1e-12)
        assert orc.interp_sample(ramp, 28.6, code) == pytest.approx(28.6, abs=1e-12)


def test_method_names_and_aliases():
    """As the reference's get_method_code (aliases sinc -> sinc8, spline ->
    cubic, lanczos -> lanczos3; unknown names and None -> linear), with this
    ABI's codes."""
    for rc, name in enumerate(METHODS):
        assert _abi.interp_code(name) == REF_TO_ABI[rc]
    assert _abi.interp_code("SINC") == _abi.INTERP_SINC8
    assert _abi.interp_code(" spline ") == _abi.INTERP_CUBIC
    assert _abi.interp_code("lanczos") == _abi.INTERP_LANCZOS3
    assert _abi.interp_code("unknown") == _abi.INTERP_LINEAR
    assert _abi.interp_code(None) == _abi.INTERP_LINEAR


# --------------------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def kernel():
    from pm_amd.kernels import HipKernel, KernelConfig

    k = HipKernel()
    k.initialize(KernelConfig())
    yield k
    k.cleanup()


def _job():
    from pm_amd.synthetic import reflector_job

    return reflector_job(seed=8, n_traces=5000, n_samples=501, n_out=20, nt=400, d_out=25.0, tile=10,
                         n_planes=6, n_diffractors=10, antialias="none", device="cuda")


@pytest.mark.gpu
@pytest.mark.parametrize("method", METHODS)
def test_migration_with_each_method_against_oracle(kernel, method):
    """Every interpolator through a plan: one tile within the FP32 bar of the
    FP64 oracle with the same method, fold exact; a method other than linear
    changes the image."""
    import torch

    from pm_amd.pipeline.executor import TileExecutor

    job = _job()
    cfg = replace(job.config, interpolation_method=method)
    ex = TileExecutor(kernel, job, config=cfg)
    ex.plan().run()
    img, fold = ex.assemble()
    torch.cuda.synchronize()
    t = max(ex.tiles, key=lambda t: int(t.rows.shape[0]))
    rows = t.rows.cpu().numpy().astype(np.int64)
    h = ex.host_traces(rows)

    def oracle(code):
        inp = orc.MigrateInputs(
            amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
            receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
            midpoint_y=h["midpoint_y"], x_axis=t.x_axis.cpu().numpy(), y_axis=t.y_axis.cpu().numpy(),
            t_axis_ms=job.t_axis_ms, vrms=t.vrms.cpu().numpy(), dt_in_ms=job.dt_ms,
            max_aperture_m=cfg.max_aperture_m, min_aperture_m=cfg.min_aperture_m, max_dip_deg=cfg.max_dip_degrees,
            taper_fraction=cfg.taper_fraction, apply_spreading=cfg.apply_spreading,
            apply_obliquity=cfg.apply_obliquity, grid_dx=cfg.grid_dx, grid_dy=cfg.grid_dy, interpolation=code)
        return orc.migrate(inp)

    o = oracle(_abi.interp_code(method))
    got = img[t.x0:t.x1, t.y0:t.y1].double().cpu().numpy()
    rel = np.linalg.norm(got - o.image) / np.linalg.norm(o.image)
    print(f"{method}: tile rel-L2 vs FP64 oracle {rel:.2e}")
    assert rel <= 1e-5
    np.testing.assert_array_equal(fold[t.x0:t.x1, t.y0:t.y1].cpu().numpy(), o.fold2d)
    if method != "linear":
        lin = oracle(_abi.INTERP_LINEAR).image
        assert np.linalg.norm(o.image - lin) > 10 * rel * np.linalg.norm(o.image)


@pytest.mark.gpu
def test_interpolation_refused_where_the_reference_has_none(kernel):
    """Methods other than linear: straight ray, no anti-aliasing, no counter."""
    from pm_amd.pipeline.executor import TileExecutor

    job = _job()
    for cfg in (replace(job.config, interpolation_method="sinc8", antialias="triangle"),
                replace(job.config, interpolation_method="cubic", kernel_type="curved_ray",
                        curved_ray_enabled=True)):
        ex = TileExecutor(kernel, job, config=cfg)
        with pytest.raises(ValueError, match="interpolation"):
            ex.plan()
    ex = TileExecutor(kernel, job, config=replace(job.config, interpolation_method="lanczos3"))
    with pytest.raises(ValueError, match="interpolation"):
        ex.run(count_effective=True)


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["nearest", "sinc8", "lanczos5"])
def test_interpolation_with_late_trace_start_and_1d_velocity(kernel, method):
    """Traces that start 20 ms after the output's t = 0 (sample positions near
    and below 0, the lower-bound decisions) with a 1-D velocity: the tile
    against the oracle.  The output (x, y = 200 .. 1175 m, 0 .. 898 ms) holds
    the diffractor at (1000 m, 1000 m, 0.8 s) and the traces (2 s) its whole
    diffraction, so the image is the focused event, not migrated noise."""
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.synthetic import point_diffractor_job

    job = point_diffractor_job(seed=3, n_traces=2000, n_samples=1001, n_out=40, nt=450, device="cuda")
    job.t_start_ms = 20.0
    cfg = replace(job.config, interpolation_method=method)
    ex = TileExecutor(kernel, job, config=cfg)
    ex.plan().run()
    img, fold = ex.assemble()
    t = ex.tiles[0]
    rows = t.rows.cpu().numpy().astype(np.int64)
    h = ex.host_traces(rows)
    inp = orc.MigrateInputs(
        amplitudes=h["amplitudes"], source_x=h["source_x"], source_y=h["source_y"],
        receiver_x=h["receiver_x"], receiver_y=h["receiver_y"], midpoint_x=h["midpoint_x"],
        midpoint_y=h["midpoint_y"], x_axis=t.x_axis.cpu().numpy(), y_axis=t.y_axis.cpu().numpy(),
        t_axis_ms=job.t_axis_ms, vrms=job.vrms_1d, dt_in_ms=job.dt_ms, t_start_in_ms=20.0,
        max_aperture_m=cfg.max_aperture_m, min_aperture_m=cfg.min_aperture_m, max_dip_deg=cfg.max_dip_degrees,
        taper_fraction=cfg.taper_fraction, apply_spreading=cfg.apply_spreading,
        apply_obliquity=cfg.apply_obliquity, grid_dx=cfg.grid_dx, grid_dy=cfg.grid_dy,
        interpolation=_abi.interp_code(method))
    o = orc.migrate(inp)
    got = img[t.x0:t.x1, t.y0:t.y1].double().cpu().numpy()
    rel = np.linalg.norm(got - o.image) / np.linalg.norm(o.image)
    print(f"{method} (trace start 20 ms): rel-L2 {rel:.2e}")
    assert rel <= 1e-5
    np.testing.assert_array_equal(fold[t.x0:t.x1, t.y0:t.y1].cpu().numpy(), o.fold2d)
