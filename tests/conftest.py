import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden_dir() -> Path:
    return GOLDEN


def golden_inputs(npz):
    """MigrateInputs from a golden migrate_*.npz (written by
    tests/golden/make_golden.py from the reference's own kernel)."""
    from oracle.oracle import MigrateInputs

    g = npz
    return MigrateInputs(
        amplitudes=g["amplitudes"], source_x=g["source_x"], source_y=g["source_y"],
        receiver_x=g["receiver_x"], receiver_y=g["receiver_y"],
        midpoint_x=g["midpoint_x"], midpoint_y=g["midpoint_y"],
        x_axis=g["x_axis"], y_axis=g["y_axis"], t_axis_ms=g["t_axis_ms"], vrms=g["vrms"],
        dt_in_ms=float(g["dt_in_ms"]), t_start_in_ms=float(g["t_start_in_ms"]),
        trace_weights=g["trace_weights"], max_aperture_m=float(g["max_aperture_m"]),
        min_aperture_m=float(g["min_aperture_m"]), max_dip_deg=float(g["max_dip_deg"]),
        taper_fraction=float(g["taper_fraction"]),
        apply_spreading=bool(g["apply_spreading"]), apply_obliquity=bool(g["apply_obliquity"]))
