"""Zarr v2 trace arrays and Parquet headers (CPU): round trips, edge chunks,
missing chunks (fill value), unsupported encodings, header validation and
group boundaries."""

from __future__ import annotations

import json

import numpy as np
import pytest

from pm_amd.io.zarr_v2 import ZarrV2Array, read_headers, write_headers, write_zarr_v2
from pm_amd.pipeline.streaming import group_bounds


def test_round_trip_and_edge_chunk(tmp_path):
    a = np.random.default_rng(0).standard_normal((2503, 77)).astype(np.float32)
    z = write_zarr_v2(tmp_path / "t.zarr", a, chunk_rows=1000, attrs={"sample_rate_ms": 2.0})
    assert z.shape == (2503, 77) and z.n_chunks == 3 and z.attrs["sample_rate_ms"] == 2.0
    np.testing.assert_array_equal(z.read_rows(0, 2503), a)
    np.testing.assert_array_equal(z.read_rows(999, 1001), a[999:1001])
    # the edge chunk file holds a full, zero-padded chunk (Zarr v2 layout)
    raw = np.fromfile(tmp_path / "t.zarr" / "2.0", dtype=np.float32).reshape(1000, 77)
    np.testing.assert_array_equal(raw[:503], a[2000:])
    assert not raw[503:].any()
    meta = json.loads((tmp_path / "t.zarr" / ".zarray").read_text())
    assert meta["chunks"] == [1000, 77] and meta["compressor"] is None and meta["dtype"] == "<f4"


def test_missing_chunk_reads_fill_value(tmp_path):
    a = np.ones((2000, 5), np.float32)
    z = write_zarr_v2(tmp_path / "t.zarr", a, chunk_rows=1000)
    (tmp_path / "t.zarr" / "1.0").unlink()
    got = z.read_rows(0, 2000)
    assert got[:1000].all() and not got[1000:].any()


@pytest.mark.parametrize("change,exc", [
    ({"compressor": {"id": "lz4"}}, NotImplementedError),
    ({"filters": [{"id": "delta"}]}, NotImplementedError),
    ({"chunks": [1000, 10]}, NotImplementedError),
    ({"dtype": "<i2"}, NotImplementedError),
    ({"order": "F"}, NotImplementedError),
    ({"zarr_format": 3}, ValueError),
])
def test_unsupported_encodings_raise(tmp_path, change, exc):
    write_zarr_v2(tmp_path / "t.zarr", np.zeros((10, 20), np.float32), chunk_rows=5)
    p = tmp_path / "t.zarr" / ".zarray"
    meta = json.loads(p.read_text())
    meta.update(change)
    p.write_text(json.dumps(meta))
    with pytest.raises(exc):
        ZarrV2Array.open(tmp_path / "t.zarr")


def test_headers_round_trip_and_validation(tmp_path):
    x = np.linspace(0, 1, 6)
    write_headers(tmp_path / "h.parquet", x, x + 1, x + 2, x + 3, extra={"BIN": [0, 0, 1, 1, 1, 2]})
    h = read_headers(tmp_path / "h.parquet")
    np.testing.assert_array_equal(h["REC_Y"], x + 3)
    assert group_bounds(h, "BIN", 6, 4) == [(0, 2), (2, 5), (5, 6)]
    assert group_bounds(h, None, 6, 4) == [(0, 4), (4, 6)]
    import pyarrow as pa
    import pyarrow.parquet as pq

    pq.write_table(pa.table({"trace_idx": np.arange(3), "SOU_X": np.zeros(3)}), str(tmp_path / "bad.parquet"))
    with pytest.raises(ValueError, match="missing required columns"):
        read_headers(tmp_path / "bad.parquet")
    h2 = dict(h)
    h2["BIN"] = np.array([0, 1, 0, 1, 1, 2])
    with pytest.raises(ValueError, match="group by group"):
        group_bounds(h2, "BIN", 6, 4)
