"""CPU tests of the drop-in boundary: the mirror of the reference's kernel
interface (pstm/kernels/base.py, factory.py, config/backends.py) and the C ABI
of the HIP library (loaded, every declared symbol exported, layout checked,
argument validation) — no device work."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from pm_amd.config import ComputeBackend, parse_backend
from pm_amd.kernels import (KernelConfig, OutputTile, VelocitySlice, create_kernel,
                            create_output_tile, create_trace_block)
from pm_amd.kernels import _abi
from pm_amd.kernels.base import MigrationKernel
from pm_amd.kernels.hip_backend import HipKernel

ROOT = Path(__file__).resolve().parent.parent


# --- mirrors tests/test_kernels.py:15-70 of the reference -------------------

def test_create_trace_block_and_offset():
    b = create_trace_block(np.zeros((10, 100), np.float32), np.zeros(10), np.zeros(10),
                           np.ones(10) * 100, np.zeros(10), sample_rate_ms=2.0)
    assert b.n_traces == 10 and b.n_samples == 100
    np.testing.assert_array_almost_equal(b.offset, 100.0)
    np.testing.assert_array_equal(b.midpoint_x, 50.0)
    assert len(b.validate()) == 0
    np.testing.assert_array_equal(b.time_axis_ms[:3], [0.0, 2.0, 4.0])


def test_trace_block_validate_catches_errors():
    b = create_trace_block(np.zeros((4, 10), np.float32), np.zeros(4), np.zeros(4),
                           np.zeros(4), np.zeros(4), sample_rate_ms=2.0)
    b.source_x = np.zeros(3)
    b.amplitudes[0, 0] = np.nan
    errs = b.validate()
    assert any("source_x length mismatch" in e for e in errs)
    assert any("NaN" in e for e in errs)


def test_create_output_tile_and_reset():
    t = create_output_tile(0, 100, 10, 0, 100, 10, 0, 1000, 2)
    assert (t.nx, t.ny, t.nt) == (11, 11, 501)
    assert t.fold.shape == (11, 11, 501) and t.fold.dtype == np.int32
    t.image[:] = 1.0
    t.fold[:] = 3
    t.reset()
    assert not t.image.any() and not t.fold.any()
    assert t.dt_ms == 2.0 and t.x_range == (0.0, 100.0)


def test_kernel_config_defaults():
    c = KernelConfig()
    assert c.max_aperture_m == 5000.0 and c.min_aperture_m == 500.0
    assert c.max_dip_degrees == 45.0 and c.taper_fraction == 0.1
    assert c.apply_spreading is True and c.apply_obliquity is True
    assert c.interpolation_method == "linear" and c.antialias == "none"
    # base.py:261-265: no factor 1/2
    assert c.get_aperture_at_time(1000.0, 2000.0) == pytest.approx(2000.0)


def test_velocity_slice():
    v = VelocitySlice(vrms=np.arange(24.0).reshape(2, 3, 4), is_1d=False)
    assert v.get_velocity_at(1, 2, 3) == 23.0
    np.testing.assert_array_equal(v.get_velocity_1d(), [16, 17, 18, 19])


def test_backend_parsing_and_factory():
    assert parse_backend("hip") is ComputeBackend.HIP
    assert parse_backend(" HIP ") is ComputeBackend.HIP
    assert parse_backend(ComputeBackend.AUTO) is ComputeBackend.AUTO
    with pytest.raises(ValueError):
        parse_backend("cuda")
    with pytest.raises(TypeError):
        parse_backend(3)
    k = create_kernel("hip")
    assert isinstance(k, HipKernel) and k.name == "hip"
    assert isinstance(k, MigrationKernel)
    assert hasattr(k, "migrate_tile_curved_ray") and hasattr(k, "migrate_tile_vti")


def test_uninitialised_kernel_raises():
    k = HipKernel()
    tb = create_trace_block(np.zeros((2, 10), np.float32), np.zeros(2), np.zeros(2),
                            np.zeros(2), np.zeros(2), sample_rate_ms=2.0)
    out = create_output_tile(0, 10, 10, 0, 10, 10, 0, 10, 2)
    with pytest.raises(RuntimeError, match="not initialized"):
        k.migrate_tile(tb, out, VelocitySlice(np.full(out.nt, 2000.0)))


def test_initialize_without_gpu_fails_loudly():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        HipKernel().initialize(KernelConfig())


# --- the C ABI ----------------------------------------------------------------

def _declared_functions() -> list[str]:
    text = (ROOT / "include" / "pstm_hip.h").read_text()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(pstm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    declared = _declared_functions()
    assert set(declared) == set(_abi.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_abi_layout_matches_header():
    lib = _abi.load_library()
    assert lib.pstm_hip_abi_version() == _abi.ABI_VERSION
    assert lib.pstm_hip_args_size() == ctypes.sizeof(_abi.PstmMigrateArgs)
    # every field of the C struct appears in the ctypes mirror, in order
    text = (ROOT / "include" / "pstm_hip.h").read_text()
    body = text[text.index("typedef struct PstmMigrateArgs {"):text.index("} PstmMigrateArgs;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S).split("{", 1)[1]
    c_fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            first, *rest = [d.strip() for d in decl.split(",")]
            c_fields += [re.findall(r"(\w+)$", first)[0], *rest]
    assert c_fields == [f[0] for f in _abi.PstmMigrateArgs._fields_]


def test_invalid_arguments_are_rejected_without_device_work():
    lib = _abi.load_library()
    assert lib.pstm_hip_migrate_tile(None, None) == 1
    assert "NULL" in _abi.last_error(lib)
    a = _abi.PstmMigrateArgs()
    a.nx = a.ny = a.nt = 0
    assert lib.pstm_hip_migrate_tile(ctypes.byref(a), None) == 1
    assert "invalid sizes" in _abi.last_error(lib)
    a.n_samples, a.nx, a.ny, a.nt = 10, 2, 2, 5
    a.x_axis = a.y_axis = a.t_axis_ms = a.vrms = a.image = a.fold = 8
    a.dt_in_ms, a.max_aperture_m = 2.0, 100.0
    a.traveltime, a.aa_mode, a.aa_max_half_width = _abi.TT_VTI, _abi.AA_TRIANGLE, 4
    assert lib.pstm_hip_migrate_tile(ctypes.byref(a), None) == 1
    assert "straight-ray" in _abi.last_error(lib)
    gs = (ctypes.c_int32 * 3)(0, 5, 3)
    assert lib.pstm_hip_migrate_bins(ctypes.byref(a), gs, 2, None) == 1
    assert "non-decreasing" in _abi.last_error(lib)
    # learned traveltime without a network, bad lattice spacing
    a.aa_mode = _abi.AA_NONE
    a.traveltime = _abi.TT_VTI_LEARNED
    assert lib.pstm_hip_migrate_tile(ctypes.byref(a), None) == 1
    assert "mlp_params" in _abi.last_error(lib)
    a.mlp_params, a.mlp_node_spacing = 8, 2
    assert lib.pstm_hip_migrate_tile(ctypes.byref(a), None) == 1
    assert "mlp_node_spacing" in _abi.last_error(lib)
    a.mlp_node_spacing, a.mlp_node_growth = 8, 2000
    assert lib.pstm_hip_migrate_tile(ctypes.byref(a), None) == 1
    assert "mlp_node_growth" in _abi.last_error(lib)
    a.mlp_node_growth, a.mlp_node_max = 0, 4
    assert lib.pstm_hip_migrate_tile(ctypes.byref(a), None) == 1
    assert "mlp_node_max" in _abi.last_error(lib)
    a.mlp_node_max = 0
    # plans, tuning knobs
    h = ctypes.c_void_p()
    assert lib.pstm_hip_plan_create(None, None, 0, ctypes.byref(h)) == 1
    assert lib.pstm_hip_plan_run(None, 1, None) == 1
    assert lib.pstm_hip_set_tuning(5, 0) == 1 and lib.pstm_hip_set_tuning(0, 17) == 1
    assert lib.pstm_hip_set_pipeline(2) == 1
    assert lib.pstm_hip_set_tuning(0, 0) == 0 and lib.pstm_hip_set_pipeline(-1) == 0
    assert lib.pstm_mlp_set_precision(3) == 1 and lib.pstm_mlp_set_precision(0) == 0


def test_roofline_constants_frozen():
    from pm_amd import roofline

    assert roofline.FLOPS_PER_CONTRIBUTION["straight"] == 19
    assert roofline.FLOPS_PER_CONTRIBUTION["straight_noweights"] == 16
    assert roofline.MI355X_FP32_VALU_TFLOPS == 157.3


def _plan_common():
    a = _abi.PstmMigrateArgs()
    a.n_traces, a.n_samples, a.nt = 4, 10, 5
    # fake, never dereferenced: the overlap test runs before any device work
    a.amplitudes = a.source_x = a.source_y = a.receiver_x = a.receiver_y = 8
    a.midpoint_x = a.midpoint_y = a.t_axis_ms = 8
    a.dt_in_ms, a.max_aperture_m = 2.0, 100.0
    return a


def _desc(image: int, fold: int, nx: int = 2, ny: int = 3):
    d = _abi.PstmTileDesc()
    d.x_axis = d.y_axis = d.vrms = 8
    d.image, d.fold = image, fold
    d.trace_index, d.n_traces, d.nx, d.ny = 8, 4, nx, ny
    return d


def test_plan_rejects_tiles_that_share_image_or_fold_memory():
    """The gather adds into image cells and the fold counts into fold cells
    without atomics, so two descriptors over one buffer would lose updates:
    plan creation refuses them before touching the device."""
    lib = _abi.load_library()
    a = _plan_common()
    img_bytes = 2 * 3 * 5 * 4          # nx ny nt f32
    h = ctypes.c_void_p()
    cases = {
        "same image": [_desc(0x10000, 0x90000), _desc(0x10000, 0xA0000)],
        "image ranges overlap by one cell": [_desc(0x10000, 0x90000),
                                             _desc(0x10000 + img_bytes - 4, 0xA0000)],
        "fold ranges overlap": [_desc(0x10000, 0x90000), _desc(0x20000, 0x90000 + 8)],
        "fold inside an image": [_desc(0x10000, 0x90000), _desc(0x20000, 0x10000 + 16)],
        "third tile overlaps the first": [_desc(0x10000, 0x90000, nx=20), _desc(0x10100, 0xA0000),
                                          _desc(0x30000, 0xB0000)],
    }
    for name, descs in cases.items():
        arr = (_abi.PstmTileDesc * len(descs))(*descs)
        rc = lib.pstm_hip_plan_create(ctypes.byref(a), arr, len(descs), ctypes.byref(h))
        assert rc == 1, name
        assert "overlaps" in _abi.last_error(lib), name
    # adjacent, disjoint ranges pass the check (and then fail only for want of
    # a device here)
    descs = [_desc(0x10000, 0x90000), _desc(0x10000 + img_bytes, 0x90000 + 2 * 3 * 4)]
    arr = (_abi.PstmTileDesc * 2)(*descs)
    rc = lib.pstm_hip_plan_create(ctypes.byref(a), arr, 2, ctypes.byref(h))
    assert "overlaps" not in _abi.last_error(lib)
    import torch

    if torch.cuda.device_count() == 0:
        assert rc != 0


def test_multiply_high_division_rule():
    """The gather divides lattice-row, node and segment indices by
    batch-dependent values with one multiply-high, m = floor((2^32 - 1) / d) + 1
    and a / d = (a m) >> 32 (div_magic / div_by in pstm_hip.hip), claimed exact
    for 0 <= a < 2^16 and 2 <= d < 2^16; the indices stay below 2^16 because
    the tables they index fit the 160 KB LDS.  Checked for every a and a range
    of divisors, the largest included."""
    a = np.arange(1 << 16, dtype=np.uint64)
    divisors = list(range(2, 1025)) + [1031, 4093, 4096, 12345, 40000, 65521, 65535]
    for d in divisors:
        m = np.uint64((2**32 - 1) // d + 1)
        assert m < 2**32
        q = (a * m) >> np.uint64(32)
        np.testing.assert_array_equal(q, a // np.uint64(d), err_msg=f"d = {d}")
    # the rule fails past its range (a = 2^16 needs more precision for some d):
    # the kernel relies on the bound, not on luck
    a_big = np.arange(1 << 16, 1 << 22, dtype=np.uint64)
    d = 65535
    m = np.uint64((2**32 - 1) // d + 1)
    assert np.any(((a_big * m) >> np.uint64(32)) != a_big // np.uint64(d))
