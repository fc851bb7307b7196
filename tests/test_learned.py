"""Learned VTI traveltime fused into the gather (PSTM_TT_VTI_LEARNED).

* Traveltime read-back: a single trace whose samples are a ramp A[i] = i,
  all weights 1, makes every image sample equal the position u = t / dt of its
  contribution, so the gather's own traveltimes can be compared with the
  bf16 NumPy/torch restatement of the lattice inference
  (pm_amd.surrogate.lattice_times over oracle-pinned bf16 numerics), and the
  exact VTI path with the f64 VTI time.
* Job level: on a small VTI diffractor job the network trained (f32) on the
  job's model and evaluated with split bf16 weights passes the per-tile
  held-out check, and the learned image is within relative L2 1e-3 of the
  exact-traveltime image, which the image the method gives without the
  network (isotropic straight-ray times in the same v_rms) misses by far.
* Fallback: a tile that fails the held-out check migrates with the exact
  traveltime, bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

from pm_amd.kernels import _abi
from pm_amd.lattice import GEOMETRIC, Lattice

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kernel():
    from pm_amd.kernels import HipKernel, KernelConfig

    k = HipKernel()
    k.initialize(KernelConfig())
    yield k
    k.cleanup()


def _ramp_case(seed: int, n_samples=1001, nt=480, nxy=4):
    import torch

    rng = np.random.default_rng(seed)
    dev = "cuda"
    sx, sy = rng.uniform(0, 1500, 2)
    rx, ry = rng.uniform(0, 1500, 2)
    xa = np.linspace(600.0, 900.0, nxy)
    ya = np.linspace(500.0, 800.0, nxy)
    t_ms = 2.0 * np.arange(nt)
    tt = t_ms[None, None, :] / 1000.0
    v = 1600.0 + 900.0 * tt * (1.0 + 0.05 * rng.standard_normal((nxy, nxy, 1)))
    eta = np.clip(0.1 + 0.05 * np.sin(3 * tt + rng.uniform(0, 3, (nxy, nxy, 1))), 0, 0.2)
    amp = np.arange(n_samples, dtype=np.float32)[None, :]
    from pm_amd.kernels.hip_backend import DeviceTraces

    f64 = torch.float64
    d = lambda a: torch.as_tensor(np.atleast_1d(np.asarray(a, dtype=np.float64)), device=dev)  # noqa: E731
    tr = DeviceTraces(torch.as_tensor(amp, device=dev), d(sx), d(sy), d(rx), d(ry), d((sx + rx) / 2),
                      d((sy + ry) / 2), None, 2.0, 0.0)
    return dict(tr=tr, xa=xa, ya=ya, t_ms=t_ms, v=v, eta=eta, src=(sx, sy), rec=(rx, ry),
                xa_d=torch.as_tensor(xa, dtype=f64, device=dev), ya_d=torch.as_tensor(ya, dtype=f64, device=dev),
                t_d=torch.as_tensor(t_ms, dtype=f64, device=dev), v_d=torch.as_tensor(v, dtype=f64, device=dev),
                eta_d=torch.as_tensor(eta, dtype=f64, device=dev))


def _ramp_config():
    from pm_amd.kernels import KernelConfig

    return KernelConfig(max_aperture_m=1e5, min_aperture_m=1e5, max_dip_degrees=89.0, taper_fraction=1e-6,
                        apply_spreading=False, apply_obliquity=False, kernel_type="anisotropic_vti",
                        vti_enabled=True)


def _readback(kernel, c, traveltime, surrogate=None, node_spacing=16):
    import torch

    nx = ny = len(c["xa"])
    img = torch.zeros((nx, ny, len(c["t_ms"])), dtype=torch.float32, device="cuda")
    fold = torch.zeros((nx, ny), dtype=torch.int32, device="cuda")
    kernel.migrate_device(c["tr"], img, fold, c["xa_d"], c["ya_d"], c["t_d"], c["v_d"], _ramp_config(),
                          traveltime=traveltime, eta=c["eta_d"], surrogate=surrogate, node_spacing=node_spacing)
    torch.cuda.synchronize()
    return img.double().cpu().numpy() * 2e-3  # u * dt = t


def _expected(c, learned, surrogate=None, node_spacing=16):
    import torch

    from pm_amd.surrogate import lattice_times, pillar_group_ends, vti_leg_times

    nx = ny = len(c["xa"])
    nt = len(c["t_ms"])
    X, Y, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nt), indexing="ij")
    pil = torch.as_tensor((X * ny + Y).ravel(), device="cuda")
    k = torch.as_tensor(K.ravel(), device="cuda")
    ox, oy = c["xa_d"][pil // ny], c["ya_d"][pil % ny]
    ds = ((ox - c["src"][0]) ** 2 + (oy - c["src"][1]) ** 2).sqrt()
    dr = ((ox - c["rec"][0]) ** 2 + (oy - c["rec"][1]) ** 2).sqrt()
    vflat, eflat = c["v_d"].reshape(-1, nt), c["eta_d"].reshape(-1, nt)
    v_at = lambda kk: vflat[pil, kk]  # noqa: E731
    e_at = lambda kk: eflat[pil, kk]  # noqa: E731
    tau_axis = c["t_d"] / 2000.0
    if learned:
        # the network at three pillars of each pillar's group of 8
        pe, w = pillar_group_ends(pil, nx * ny)
        ex_, ey_ = c["xa_d"][pe // ny], c["ya_d"][pe % ny]
        ends = {"ds": ((ex_ - c["src"][0]) ** 2 + (ey_ - c["src"][1]) ** 2).sqrt(),
                "dr": ((ex_ - c["rec"][0]) ** 2 + (ey_ - c["rec"][1]) ** 2).sqrt(),
                "v_at": lambda kk, e: vflat[pe[:, e], kk], "eta_at": lambda kk, e: eflat[pe[:, e], kk], "w": w}
        t = lattice_times(surrogate, tau_axis, k, ds, dr, v_at, e_at, node_spacing, ends)
    else:
        tau = tau_axis[k]
        t = vti_leg_times(tau, ds, v_at(k), e_at(k)) + vti_leg_times(tau, dr, v_at(k), e_at(k))
    return t.cpu().numpy().reshape(nx, ny, nt)


def _compare(got, exp, tol_samples):
    u = exp / 2e-3
    inside = (u > 1.0) & (u < 999.0 - 1.0)
    assert inside.sum() > 1000
    err = np.abs(got - exp)[inside] / 2e-3
    print("read-back error (samples): max", err.max(), "mean", err.mean())
    assert err.max() < tol_samples


def test_exact_vti_readback(kernel):
    c = _ramp_case(1)
    _compare(_readback(kernel, c, _abi.TT_VTI), _expected(c, False), 2e-3)


@pytest.mark.parametrize("spacing,terms", [(8, 1), (16, 1), (64, 1), (16, 2), (12, 1),
                                           (Lattice(8, 77, 64), 2), (Lattice(4, 256, 1024), 1)])
def test_learned_readback(kernel, spacing, terms):
    """The fused gather's learned times equal the restated lattice inference
    (an untrained network: the check is of the arithmetic, not of accuracy),
    with weights rounded to bf16 and with split (two-term) weights."""
    import torch

    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate

    norm = Normalizer(tau_max=1.0, d_max=2000.0, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
    sur = TraveltimeSurrogate.create(norm, seed=3)
    sur.weight_terms = terms
    # non-zero biases and larger outputs so every layer matters
    g = torch.Generator(device="cuda").manual_seed(5)
    sur.params += 0.05 * torch.randn(sur.params.shape, device="cuda", generator=g)
    c = _ramp_case(2)
    got = _readback(kernel, c, _abi.TT_VTI_LEARNED, sur, spacing)
    exp = _expected(c, True, sur, spacing)
    _compare(got, exp, 5e-3)


def test_learned_rejects_bad_arguments(kernel):
    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate

    c = _ramp_case(3)
    sur = TraveltimeSurrogate.create(Normalizer(1.0, 2000.0, 1500.0, 4500.0, 0.2))
    with pytest.raises(ValueError, match="node spacing"):
        _readback(kernel, c, _abi.TT_VTI_LEARNED, sur, node_spacing=2)
    with pytest.raises(ValueError, match="surrogate"):
        _readback(kernel, c, _abi.TT_VTI_LEARNED, None)


@pytest.fixture(scope="module")
def vti_small(kernel):
    """A small VTI diffractor job, its executor and a surrogate trained on its
    model."""
    import torch

    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.surrogate import Normalizer, TraveltimeSurrogate, make_training_set
    from pm_amd.synthetic import vti_diffractor_job

    job = vti_diffractor_job(n_traces=30_000, n_samples=1001, n_out=64, nt=480, tile=32, n_diffractors=40,
                             device="cuda")
    ex = TileExecutor(kernel, job)
    vc, ec = job.cube, job.meta["eta_cube"]
    d_max = job.config.max_aperture_m + 1500.0
    norm = Normalizer(tau_max=0.5, d_max=d_max, v_lo=1500.0, v_hi=4500.0, eta_max=0.2)
    tr = make_training_set(vc, ec, 2_000_000, d_max, norm, seed=1, t_max_ms=1000.0)
    # f32 training, inference with split (two-term bf16) weights
    sur = TraveltimeSurrogate.create(norm, seed=0)
    sur.weight_terms = 2
    sur.train(tr, steps=60000, batch=32768, qat_fraction=0.0)
    torch.cuda.synchronize()
    return job, ex, sur


def _image(ex, **kw):
    import torch

    ex.reset()
    ex.run(**kw)
    img, _ = ex.assemble()
    torch.cuda.synchronize()
    return img.double()


def test_learned_job_against_exact(kernel, vti_small):
    import time

    import torch

    job, ex, sur = vti_small
    errs = ex.check_surrogate(sur, node_spacing=GEOMETRIC)
    print("per-tile held-out max error (samples)", errs)
    assert all(e <= 0.25 for e in errs)
    exact = _image(ex)
    learned = _image(ex, surrogate=sur, node_spacing=GEOMETRIC)
    rel = float(torch.linalg.norm(learned - exact) / torch.linalg.norm(exact))
    # without the network: isotropic straight-ray times in the same v_rms
    from dataclasses import replace

    cfg_iso = replace(job.config, kernel_type="straight_ray", vti_enabled=False)
    saved = ex.config
    ex.config = cfg_iso
    try:
        iso = _image(ex)
    finally:
        ex.config = saved
    rel_iso = float(torch.linalg.norm(iso - exact) / torch.linalg.norm(exact))
    # speed: the same pass with each traveltime
    times = {}
    for name, kw in (("exact", {}), ("learned", {"surrogate": sur, "node_spacing": GEOMETRIC})):
        ex.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ex.run(**kw)
        torch.cuda.synchronize()
        times[name] = time.perf_counter() - t0
    print(f"image relL2 learned vs exact {rel:.3e}; without the network {rel_iso:.3e}; "
          f"pass time exact {times['exact'] * 1e3:.1f} ms, learned {times['learned'] * 1e3:.1f} ms")
    assert rel <= 1e-3
    assert rel_iso > 100 * rel


def test_failed_tile_falls_back_to_exact(kernel, vti_small):
    import torch

    job, ex, sur = vti_small
    ex.check_surrogate(sur, node_spacing=GEOMETRIC, bar=1e-9)   # no tile passes
    assert not any(t.learned_ok for t in ex.tiles)
    exact = _image(ex)
    fallback = _image(ex, surrogate=sur, node_spacing=GEOMETRIC)
    assert torch.equal(exact, fallback)


def test_learned_launch_shapes_agree(kernel, vti_small):
    """The learned gather under different launch shapes (traces per batch:
    the pair table built by wave 0 in one or two passes, the next batch's
    pair inputs staged ahead for up to 16 traces; sample blocks per wave)
    gives the same image: within one MW a shape changes only how the f32
    per-batch sums are grouped (every network row, coefficient and time is
    the same), so the images agree to f32 summation order."""
    import torch

    job, ex, sur = vti_small
    lib = _abi.load_library()
    saved = [t.learned_ok for t in ex.tiles]
    for t in ex.tiles:
        t.learned_ok = True
    try:
        for mw, caps in ((3, (16, 9, 4, 1)), (1, (16, 12, 3))):
            ref = None
            batches = set()
            for cap in caps:
                assert lib.pstm_hip_set_tuning(mw, cap) == 0
                plan = ex.plan(traveltime=_abi.TT_VTI_LEARNED, surrogate=sur, node_spacing=GEOMETRIC)
                info = plan.info()
                ex.reset()
                plan.run(use_graph=False)
                img, _ = ex.assemble()
                plan.destroy()
                img = img.double()
                batches.add(info["batch"])
                if ref is None:
                    ref = img
                    continue
                rel = float(torch.linalg.norm(img - ref) / torch.linalg.norm(ref))
                print(f"mw<={mw} batch<={cap} -> {info}: rel-L2 against the first shape {rel:.2e}")
                assert rel <= 1e-5, (mw, cap)
            assert len(batches) >= 2, batches
    finally:
        lib.pstm_hip_set_tuning(0, 0)
        for t, ok in zip(ex.tiles, saved):
            t.learned_ok = ok
        ex.reset()
