"""Reading the reference's own trace files: Blosc (zstd, bit-shuffle) Zarr v2
chunks, as pstm/data/zarr_reader.py:416-436 writes them and
ZarrTraceReader.get_traces (:237-301) reads them.

* The shuffles against explicit bit / byte loops on small arrays.
* Hand-built Blosc chunks (header, block starts, zstd / zlib / stored
  streams, split and unsplit blocks, version-2 and version-3 handling of a
  block whose element count is not a multiple of 8) decode to their input.
* The reference's benchmark input (benchmark_output/input_data): chunk 13
  of traces.zarr is committed under tests/golden/reference_input with the
  headers and metadata; every one of its 448 traces peaks within one sample
  of the time the reference's generator gives its diffractor
  (pstm/synthetic/common_offset_gathers.py:294-322, 377-395).  Where the
  reference checkout is present, all 14 chunks are decoded (13,448 x 1001,
  finite) and migrated by the oracle: the image focuses at the diffractor's
  (x, y) = (1000, 1000) m and t0 = 800 ms.  The generator puts the two-way
  zero-offset time t0 = 2 z / v = 0.4 s into each one-way leg
  (common_offset_gathers.py:319-320, 378), where the migration's DSR time
  uses t0 / 2 per leg (numba_cpu_optimized.py:189-194), so the data migrate
  to t0 = 2 x 0.4 s.
"""

from __future__ import annotations

import json
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from pm_amd.io import blosc
from pm_amd.io.zarr_v2 import ZarrV2Array, read_headers

FIX = Path(__file__).resolve().parent / "golden" / "reference_input"
REF = Path("/root/reference/benchmark_output/input_data")


def _bitshuffle_loops(data: np.ndarray, typesize: int) -> np.ndarray:
    """Forward bit-shuffle by explicit loops: for byte j and bit k, the bits
    of elements 0, 1, 2, ... packed eight to a byte, least significant first."""
    raw = data.view(np.uint8).reshape(-1, typesize)
    n = raw.shape[0]
    out = np.zeros(typesize * n, np.uint8)
    pos = 0
    for j in range(typesize):
        for k in range(8):
            for e in range(n):
                if (raw[e, j] >> k) & 1:
                    out[pos + e // 8] |= np.uint8(1 << (e % 8))
            pos += n // 8
    return out


def test_bitunshuffle_inverts_the_loop_definition():
    rng = np.random.default_rng(0)
    for typesize, n in ((4, 64), (8, 40), (2, 16), (1, 24)):
        data = rng.integers(0, 256, typesize * n, dtype=np.uint8)
        sh = _bitshuffle_loops(data, typesize)
        np.testing.assert_array_equal(blosc.bitunshuffle(sh, typesize, n), data)


def test_byteunshuffle_inverts_the_loop_definition():
    rng = np.random.default_rng(1)
    data = rng.integers(0, 256, 4 * 37 + 3, dtype=np.uint8)   # 3 trailing bytes stay in place
    sh = np.concatenate([np.array([data[e * 4 + j] for j in range(4) for e in range(37)], np.uint8), data[-3:]])
    np.testing.assert_array_equal(blosc.byteunshuffle(sh, 4, len(sh)), data)


def _chunk(data: np.ndarray, typesize: int, blocksize: int, codec: int, shuffle: str, split: bool,
           version: int = 2, store_raw_block: int = -1) -> bytes:
    """A Blosc chunk assembled from the format's definition."""
    import pyarrow as pa

    raw = data.view(np.uint8)
    nbytes = raw.size
    nblocks = (nbytes + blocksize - 1) // blocksize
    flags = (codec << 5) | (0 if split else 0x10)
    flags |= {"none": 0, "byte": 0x1, "bit": 0x4}[shuffle]
    body = bytearray()
    starts = []
    head = 16 + 4 * nblocks
    for i in range(nblocks):
        blk = raw[i * blocksize:(i + 1) * blocksize]
        n = blk.size // typesize
        if shuffle == "bit":
            if version <= 2:
                sh = _bitshuffle_loops(blk, typesize) if n % 8 == 0 else blk.copy()
            else:
                m = n - n % 8
                sh = np.concatenate([_bitshuffle_loops(blk[:m * typesize], typesize) if m else blk[:0],
                                     blk[m * typesize:]])
        elif shuffle == "byte":
            sh = np.concatenate([blk[:n * typesize].reshape(n, typesize).T.reshape(-1), blk[n * typesize:]])
        else:
            sh = blk
        nsplit = typesize if split else 1
        starts.append(head + len(body))
        for s in range(nsplit):
            part = sh[s * (blk.size // nsplit):(s + 1) * (blk.size // nsplit)].tobytes()
            if i == store_raw_block:
                comp = part
            elif codec == 4:
                comp = pa.compress(part, codec="zstd", asbytes=True)
            else:
                comp = zlib.compress(part, 3)
            body += struct.pack("<i", len(comp)) + comp
    cbytes = head + len(body)
    hdr = bytes([version, 1, flags, typesize]) + struct.pack("<III", nbytes, blocksize, cbytes)
    return hdr + struct.pack(f"<{nblocks}i", *starts) + bytes(body)


@pytest.mark.parametrize("codec,shuffle,split,version", [
    (4, "bit", False, 2), (4, "bit", False, 3), (4, "byte", True, 2), (3, "none", False, 2),
    (4, "bit", True, 2), (3, "byte", False, 3)])
def test_hand_built_chunks_decode(codec, shuffle, split, version):
    rng = np.random.default_rng(2)
    # 1,003 floats: the last block (4,012 - 2 x 1,536 = 940 bytes = 235
    # elements) is not a multiple of 8 elements
    data = (rng.standard_normal(1003) * np.linspace(0, 1, 1003)).astype(np.float32)
    buf = _chunk(data, 4, 1536, codec, shuffle, split, version, store_raw_block=1)
    out = blosc.decompress(buf)
    np.testing.assert_array_equal(out.view(np.float32), data)
    h = blosc.parse_header(buf)
    assert h.nbytes == data.nbytes and h.codec == ("zstd" if codec == 4 else "zlib")


def test_memcpyed_and_truncated_chunks():
    data = np.arange(50, dtype=np.float32)
    buf = bytes([2, 1, 0x2 | 0x10, 4]) + struct.pack("<III", 200, 200, 216) + data.tobytes()
    np.testing.assert_array_equal(blosc.decompress(buf).view(np.float32), data)
    with pytest.raises(ValueError):
        blosc.decompress(buf[:100])
    with pytest.raises(ValueError):
        blosc.parse_header(b"\x07" + buf[1:])


def _generator_times(h: dict, meta: dict) -> np.ndarray:
    """The reference generator's diffractor time: t0 = 2 z / v
    (common_offset_gathers.py:378) in each leg of
    sqrt(t0^2 + (d / v)^2) (:316-321)."""
    d = meta["diffractors"][0]
    v = meta["velocity_ms"]
    t0 = 2.0 * d["z"] / v
    ds = np.hypot(h["SOU_X"] - d["x"], h["SOU_Y"] - d["y"])
    dr = np.hypot(h["REC_X"] - d["x"], h["REC_Y"] - d["y"])
    return np.sqrt(t0 ** 2 + (ds / v) ** 2) + np.sqrt(t0 ** 2 + (dr / v) ** 2)


def test_committed_reference_chunk_decodes_to_the_generated_events():
    arr = ZarrV2Array.open(FIX / "traces.zarr")
    assert arr.shape == (13448, 1001) and arr.compressor["cname"] == "zstd"
    tr = arr.read_rows(13000, 13448)
    assert tr.shape == (448, 1001) and np.isfinite(tr).all()
    meta = json.loads((FIX / "metadata.json").read_text())
    h = {k: v[13000:13448] for k, v in read_headers(FIX / "headers.parquet").items()}
    t = _generator_times(h, meta)
    peak = np.argmax(np.abs(tr), axis=1) * meta["dt_ms"] / 1000.0
    assert np.max(np.abs(peak - t)) <= meta["dt_ms"] / 1000.0 + 1e-9
    assert 0.9 < np.abs(tr).max() < 1.1


@pytest.mark.skipif(not REF.exists(), reason="the reference checkout is not present")
def test_reference_input_decodes_and_migrates_to_the_diffractor():
    from oracle import oracle as orc

    arr = ZarrV2Array.open(REF / "traces.zarr")
    tr = arr.read_rows(0, arr.n_traces)
    assert tr.shape == (13448, 1001) and tr.dtype == np.float32 and np.isfinite(tr).all()
    meta = json.loads((REF / "metadata.json").read_text())
    h = read_headers(REF / "headers.parquet")
    peak = np.argmax(np.abs(tr), axis=1) * 0.002
    assert np.max(np.abs(peak - _generator_times(h, meta))) <= 0.002 + 1e-9
    ax = 1000.0 + 25.0 * np.arange(-4, 5)
    inp = orc.MigrateInputs(
        amplitudes=tr, source_x=h["SOU_X"], source_y=h["SOU_Y"], receiver_x=h["REC_X"], receiver_y=h["REC_Y"],
        midpoint_x=(h["SOU_X"] + h["REC_X"]) / 2, midpoint_y=(h["SOU_Y"] + h["REC_Y"]) / 2, x_axis=ax,
        y_axis=ax.copy(), t_axis_ms=2.0 * np.arange(600), vrms=np.full(600, meta["velocity_ms"]),
        dt_in_ms=2.0, max_aperture_m=1500.0, min_aperture_m=1500.0)
    res = orc.migrate(inp)
    ix, iy, k = np.unravel_index(np.argmax(np.abs(res.image)), res.image.shape)
    assert (ax[ix], ax[iy]) == (1000.0, 1000.0)
    assert abs(2.0 * k - 800.0) <= 2.0


@pytest.mark.gpu
def test_reference_chunk_migrates_on_the_gpu_like_the_oracle():
    """The 448 decoded traces of the committed reference chunk through the
    drop-in migrate_tile of the hip backend against the FP64 oracle: fold
    bit-exact, image relative L2 <= 1e-5."""
    from oracle import oracle as orc
    from pm_amd.kernels import KernelConfig, VelocitySlice, create_kernel, create_output_tile, create_trace_block

    arr = ZarrV2Array.open(FIX / "traces.zarr")
    tr = arr.read_rows(13000, 13448)
    h = {k: v[13000:13448] for k, v in read_headers(FIX / "headers.parquet").items()}
    meta = json.loads((FIX / "metadata.json").read_text())
    v = float(meta["velocity_ms"])
    print(f"[reference chunk] {tr.shape[0]} traces x {tr.shape[1]}, v {v}", flush=True)
    out = create_output_tile(800.0, 1200.0, 25.0, 800.0, 1200.0, 25.0, 0.0, 1198.0, 2.0)
    cfg = KernelConfig(max_aperture_m=1000.0, min_aperture_m=300.0, max_dip_degrees=50.0)
    k = create_kernel("hip")
    k.initialize(cfg)
    try:
        tb = create_trace_block(tr, h["SOU_X"], h["SOU_Y"], h["REC_X"], h["REC_Y"], sample_rate_ms=2.0)
        k.migrate_tile(tb, out, VelocitySlice(np.full(out.nt, v)), cfg)
    finally:
        k.cleanup()
    inp = orc.MigrateInputs(
        amplitudes=tr, source_x=h["SOU_X"], source_y=h["SOU_Y"], receiver_x=h["REC_X"], receiver_y=h["REC_Y"],
        midpoint_x=(h["SOU_X"] + h["REC_X"]) / 2, midpoint_y=(h["SOU_Y"] + h["REC_Y"]) / 2,
        x_axis=out.x_axis, y_axis=out.y_axis, t_axis_ms=out.t_axis_ms, vrms=np.full(out.nt, v), dt_in_ms=2.0,
        max_aperture_m=1000.0, min_aperture_m=300.0, max_dip_deg=50.0)
    ref = orc.migrate(inp)
    np.testing.assert_array_equal(out.fold[:, :, 0], ref.fold2d)
    rel = np.linalg.norm(out.image - ref.image) / np.linalg.norm(ref.image)
    print(f"[reference chunk] image rel-L2 vs FP64 oracle {rel:.3e}, fold sum {int(ref.fold2d.sum())}")
    assert ref.fold2d.sum() > 0 and rel <= 1e-5
