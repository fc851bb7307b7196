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


def test_subset_is_the_central_tiles_and_roofline_arithmetic():
    """The timed step's tile subset is fixed: the tiles nearest the output
    grid's centre; the roofline entry is F x E over the launch time against
    the FP32 VALU peak."""
    from types import SimpleNamespace

    mod = _bench_module()
    tiles = [SimpleNamespace(x0=x, x1=x + 32, y0=y, y1=y + 32) for x in range(0, 512, 32) for y in range(0, 512, 32)]
    ex = SimpleNamespace(shape=(512, 512, 1500), tiles=tiles)
    four = mod.subset_tiles(ex, 4)
    assert sorted((tiles[i].x0, tiles[i].y0) for i in four) == [(224, 224), (224, 256), (256, 224), (256, 256)]
    assert mod.subset_tiles(ex, 1) == mod.subset_tiles(ex, 1)
    rl = mod.roofline_entry(29, 10**12, 1000.0, "k")
    assert rl["achieved"] == pytest.approx(29.0) and rl["frac"] == pytest.approx(29.0 / 157.3, abs=1e-4)
    assert rl["bound"] == "valu" and rl["unit"] == "TFLOP/s"
