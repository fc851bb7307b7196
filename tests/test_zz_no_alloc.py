"""GPU test of the allocation-free entry point pstm_hip_migrate_tile_noalloc:
called through the C ABI, its image, fold and E equal those of
pstm_hip_migrate_tile on the same inputs (the same kernels on the same data:
bit for bit), on two calls in a row, the second with a smaller input (the
fixed scratch keeps the first call's contents), with 3-D velocity and with
triangle anti-aliasing; a call whose temporaries exceed the scratch is
rejected with status 1 and changes nothing."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from pm_amd.kernels import KernelConfig, create_kernel
from pm_amd.kernels import _abi
from pm_amd.kernels.hip_backend import DeviceTraces

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kernel():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    k = create_kernel("hip")
    k.initialize(KernelConfig())
    yield k
    k.cleanup()


def _case(seed, ntr, ns, nx, ny, nt, vel3d):
    import torch

    rng = np.random.default_rng(seed)
    mx = rng.uniform(-200, 1200, ntr)
    my = rng.uniform(-200, 1000, ntr)
    h = rng.uniform(0, 1500, ntr)
    az = rng.uniform(0, 2 * np.pi, ntr)
    sx, sy = mx - h / 2 * np.cos(az), my - h / 2 * np.sin(az)
    rx, ry = mx + h / 2 * np.cos(az), my + h / 2 * np.sin(az)
    d = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    amp = torch.as_tensor(rng.standard_normal((ntr, ns)).astype(np.float32), device="cuda")
    dt = DeviceTraces(amp, d(sx), d(sy), d(rx), d(ry), d(mx), d(my), d(rng.uniform(0.5, 1.5, ntr)), 2.0, 0.0)
    t = 4.0 * np.arange(nt)
    v = 1600 + 500 * t / 1000
    if vel3d:
        v = v[None, None, :] * (1 + 0.1 * np.linspace(0, 1, nx)[:, None, None]
                                - 0.05 * np.linspace(0, 1, ny)[None, :, None])
    return dt, d(20.0 * np.arange(nx) + 100), d(20.0 * np.arange(ny) + 80), d(t), d(v)


def _call(kernel, fn, dt, xa, ya, ta, v, cfg, image0, fold0):
    import torch

    image, fold = image0.clone(), fold0.clone()
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    a, keep = kernel._args(dt, image, fold, xa, ya, ta, v, cfg, None, None, None, counter, None, 16)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(ctypes.byref(a), stream)
    torch.cuda.synchronize()
    del keep
    return rc, image, fold, int(counter.item())


@pytest.mark.parametrize("aa", ["none", "triangle"])
def test_noalloc_equals_migrate_tile_on_two_calls(kernel, aa):
    import torch

    lib = _abi.load_library()
    lib.pstm_hip_profile_reset(0)
    cfg = KernelConfig(max_aperture_m=900.0, min_aperture_m=200.0, max_dip_degrees=50.0, taper_fraction=0.15,
                       grid_dx=20.0, grid_dy=20.0, antialias=aa, aa_max_half_width=8)
    # a large call first, then a smaller one (scratch keeps the first call's
    # contents; nothing of it may leak into the second)
    for seed, ntr, ns, nx, ny, nt, vel3d in [(1, 3000, 700, 24, 20, 300, True), (2, 700, 500, 9, 7, 200, False)]:
        dt, xa, ya, ta, v = _case(seed, ntr, ns, nx, ny, nt, vel3d)
        img0 = torch.as_tensor(np.random.default_rng(seed).standard_normal((nx, ny, nt)).astype(np.float32),
                               device="cuda")
        fold0 = torch.full((nx, ny), 3, dtype=torch.int32, device="cuda")
        print(f"[no-alloc] aa={aa} seed={seed} traces={ntr} tile={nx}x{ny}x{nt} vel3d={vel3d}", flush=True)
        rc1, im1, f1, e1 = _call(kernel, lib.pstm_hip_migrate_tile, dt, xa, ya, ta, v, cfg, img0, fold0)
        rc2, im2, f2, e2 = _call(kernel, lib.pstm_hip_migrate_tile_noalloc, dt, xa, ya, ta, v, cfg, img0, fold0)
        assert rc1 == 0 and rc2 == 0, _abi.last_error(lib)
        assert torch.equal(f1, f2)
        assert e1 == e2 and e1 > 0
        assert torch.equal(im1, im2)
        assert float((im1 - img0).abs().max()) > 0


def test_noalloc_rejects_calls_beyond_the_scratch(kernel):
    import torch

    lib = _abi.load_library()
    lib.pstm_hip_profile_reset(0)
    cfg = KernelConfig(max_aperture_m=900.0, min_aperture_m=200.0)
    # 3-D velocity on 128 x 128 pillars x 1100 samples: 16 B per (pillar,
    # sample) of per-sample terms alone is 288 MB > 64 MiB
    dt, xa, ya, ta, v = _case(3, 50, 300, 128, 128, 1100, True)
    img0 = torch.zeros((128, 128, 1100), dtype=torch.float32, device="cuda")
    fold0 = torch.zeros((128, 128), dtype=torch.int32, device="cuda")
    rc, im, f, e = _call(kernel, lib.pstm_hip_migrate_tile_noalloc, dt, xa, ya, ta, v, cfg, img0, fold0)
    assert rc == 1
    assert "allocation-free" in _abi.last_error(lib)
    assert not bool(im.any()) and not bool(f.any()) and e == 0
