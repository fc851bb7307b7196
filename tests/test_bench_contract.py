"""The parts of bench.py's contract that need no GPU: one GPU only (a
request for more exits with a message), and the HBM traffic it reports for
the reflector gather is the one the committed rocprofv3 summary derives from
its FETCH_SIZE / WRITE_SIZE counters and calibration kernel."""

from __future__ import annotations

import importlib.util
import re
import subprocess
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
PROFILE = REPO / "profiles" / "k_migrate_reflectors_aa.md"


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


def _row(text: str, label: str) -> tuple[float, float]:
    line = next(l for l in text.splitlines() if l.startswith("| ") and label in l)
    cells = [c.strip() for c in line.strip().strip("|").split("|")]
    return float(cells[1].replace(",", "")), float(cells[2].replace(",", ""))


def test_reported_traffic_matches_the_profile_counters():
    if not PROFILE.exists():
        pytest.skip("profile summary not present")
    text = PROFILE.read_text()
    cal_fetch, cal_write = _row(text, "calibration `bitwise_not`")
    mig_fetch, mig_write = _row(text, "`k_migrate` (plan, all tiles)")
    # the calibration kernel reads and writes 256 MiB = 262,144 KiB each
    m = re.search(r"`bitwise_not`, ([\d,]+) KiB read and written", text)
    true_kib = float(m.group(1).replace(",", ""))
    fetch_scale = round(true_kib / cal_fetch, 3)
    write_scale = round(true_kib / cal_write, 3)
    assert fetch_scale == pytest.approx(2.0) and write_scale == pytest.approx(1.0)
    expected = (mig_fetch * fetch_scale + mig_write * write_scale) * 1024
    traffic = _bench_module().TRAFFIC["reflectors"]
    assert traffic == pytest.approx(expected, rel=1e-9)
    # and the summary states the same total in TB
    tb = float(re.search(r"the gather fetches ([\d.]+) TB", text).group(1))
    assert mig_fetch * fetch_scale * 1024 / 1e12 == pytest.approx(tb, abs=0.01)
