"""The parts of bench.py's contract that need no GPU: one GPU only (a
request for more exits with a message), the fixed tile subset a step
migrates, and the roofline arithmetic."""

from __future__ import annotations

import importlib.util
import subprocess
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent


def _bench_module():
    spec = importlib.util.spec_from_file_location("bench_contract_probe", REPO / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr
    assert p.stdout.strip() == ""


def test_region_is_central_and_tile_aligned_and_roofline_arithmetic():
    """The timed step's region is fixed: the central pillars, aligned so that
    tiles of every tuned edge cover it exactly; the roofline entry is F x E
    over the launch time against the FP32 VALU peak."""
    from types import SimpleNamespace

    from pm_amd.autotune import region_tiles

    mod = _bench_module()
    x0, x1, y0, y1 = mod.central_region(512, 512, 32, 32)
    assert (x0, x1, y0, y1) == (224, 256, 224, 256)
    for edge, count in ((16, 4), (32, 1)):
        tiles = [SimpleNamespace(x0=x, x1=x + edge, y0=y, y1=y + edge) for x in range(0, 512, edge)
                 for y in range(0, 512, edge)]
        assert len(region_tiles(SimpleNamespace(tiles=tiles), x0, x1, y0, y1)) == count
    rl = mod.roofline_entry(29, 10**12, 1000.0, "k")
    assert rl["achieved"] == pytest.approx(29.0) and rl["frac"] == pytest.approx(29.0 / 157.3, abs=1e-4)
    assert rl["bound"] == "valu" and rl["unit"] == "TFLOP/s"
