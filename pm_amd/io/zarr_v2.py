"""Minimal Zarr v2 array access for trace data, no zarr package needed.

The reference stores traces as a 2-D Zarr array (n_traces, n_samples),
float32, chunked by rows, with sample_rate_ms / start_time_ms attributes
(pstm/data/zarr_reader.py:56-301; chunks of (1000, n_samples),
:416-436).  The on-disk format is Zarr v2: a ``.zarray`` JSON document
(shape, chunks, dtype, compressor, fill_value, order, filters), an optional
``.zattrs`` JSON document, and one file per chunk named ``"<i>.<j>"``
holding the chunk's C-order bytes (full chunk shape, the edge chunk padded).
This module writes the uncompressed form (``compressor: null``) and reads
it and the Blosc-compressed form the reference writes (numcodecs Blosc,
zstd, bit-shuffle: pm_amd.io.blosc).
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_DTYPES = {"<f4": np.float32, "<f8": np.float64}


@dataclass
class ZarrV2Array:
    """A 2-D row-chunked Zarr v2 array on disk (traces x samples)."""

    path: Path
    shape: tuple[int, int]
    chunk_rows: int
    dtype: np.dtype
    attrs: dict
    compressor: dict | None = None   # numcodecs-style {"id": "blosc", ...}, or None

    @classmethod
    def open(cls, path) -> "ZarrV2Array":
        path = Path(path)
        meta_path = path / ".zarray"
        if not meta_path.exists():
            raise FileNotFoundError(f"Zarr array not found: {path}")
        meta = json.loads(meta_path.read_text())
        if meta.get("zarr_format") != 2:
            raise ValueError(f"{path}: zarr_format {meta.get('zarr_format')} (expected 2)")
        comp = meta.get("compressor")
        if meta.get("filters"):
            raise NotImplementedError(f"{path}: filters {meta.get('filters')} are not supported")
        if comp is not None and comp.get("id") != "blosc":
            raise NotImplementedError(f"{path}: compressor {comp.get('id')!r}; uncompressed or blosc chunks only")
        if meta.get("order", "C") != "C":
            raise NotImplementedError(f"{path}: order {meta.get('order')} (expected C)")
        dt = meta.get("dtype")
        if dt not in _DTYPES:
            raise NotImplementedError(f"{path}: dtype {dt} (expected <f4 or <f8)")
        shape = tuple(int(s) for s in meta["shape"])
        chunks = tuple(int(c) for c in meta["chunks"])
        if len(shape) != 2 or len(chunks) != 2:
            raise ValueError(f"{path}: expected a 2-D array, got shape {shape}")
        if chunks[1] != shape[1]:
            raise NotImplementedError(f"{path}: chunks {chunks} split the sample axis; expected whole rows")
        if meta.get("dimension_separator", ".") != ".":
            raise NotImplementedError(f"{path}: dimension_separator {meta['dimension_separator']!r}")
        attrs_path = path / ".zattrs"
        attrs = json.loads(attrs_path.read_text()) if attrs_path.exists() else {}
        return cls(path, shape, chunks[0], np.dtype(_DTYPES[dt]), attrs, comp)

    @property
    def n_traces(self) -> int:
        return self.shape[0]

    @property
    def n_samples(self) -> int:
        return self.shape[1]

    @property
    def n_chunks(self) -> int:
        return (self.shape[0] + self.chunk_rows - 1) // self.chunk_rows

    def chunk_range(self, i: int) -> tuple[int, int]:
        r0 = i * self.chunk_rows
        return r0, min(r0 + self.chunk_rows, self.shape[0])

    def chunk_file(self, i: int) -> Path:
        return self.path / f"{i}.0"

    def read_chunk_into(self, i: int, out: np.ndarray) -> int:
        """Read the valid rows of chunk i into ``out`` (C-contiguous, at least
        rows x n_samples of this array's dtype; a missing chunk file is the
        fill value 0).  Returns the number of rows.  Releases the GIL while
        reading, so several threads read in parallel."""
        r0, r1 = self.chunk_range(i)
        n = (r1 - r0) * self.shape[1]
        flat = out.reshape(-1)[:n]
        if flat.dtype != self.dtype or not flat.flags.c_contiguous:
            raise ValueError("output buffer must be C-contiguous with the array's dtype")
        f = self.chunk_file(i)
        if not f.exists():
            flat[:] = 0
            return r1 - r0
        if self.compressor is not None:
            from pm_amd.io import blosc

            raw = blosc.decompress(f.read_bytes())
            full = self.chunk_rows * self.shape[1] * self.dtype.itemsize
            if raw.nbytes != full:
                raise IOError(f"{f}: decoded chunk holds {raw.nbytes} bytes, expected {full}")
            flat.view(np.uint8)[:] = raw[:flat.nbytes]
            return r1 - r0
        view = memoryview(flat.view(np.uint8))
        with open(f, "rb", buffering=0) as fh:
            got = 0
            while got < view.nbytes:
                k = fh.readinto(view[got:])
                if not k:
                    raise IOError(f"{f}: short chunk ({got} of {view.nbytes} bytes)")
                got += k
        return r1 - r0

    def read_rows(self, r0: int, r1: int) -> np.ndarray:
        """Rows [r0, r1) as a new array (host; for tests and small reads)."""
        out = np.empty((r1 - r0, self.shape[1]), self.dtype)
        buf = np.empty((self.chunk_rows, self.shape[1]), self.dtype)
        for i in range(r0 // self.chunk_rows, (r1 - 1) // self.chunk_rows + 1 if r1 > r0 else 0):
            c0, c1 = self.chunk_range(i)
            self.read_chunk_into(i, buf)
            a, b = max(r0, c0), min(r1, c1)
            out[a - r0:b - r0] = buf[a - c0:b - c0]
        return out


def write_zarr_v2(path, data, chunk_rows: int = 1000, attrs: dict | None = None) -> ZarrV2Array:
    """Write a 2-D float32 array (NumPy, or a torch tensor on any device,
    written chunk by chunk) as an uncompressed Zarr v2 array."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    n, ns = (int(s) for s in data.shape)
    meta = {"zarr_format": 2, "shape": [n, ns], "chunks": [chunk_rows, ns], "dtype": "<f4",
            "compressor": None, "fill_value": 0.0, "order": "C", "filters": None}
    (path / ".zarray").write_text(json.dumps(meta, indent=1))
    if attrs:
        (path / ".zattrs").write_text(json.dumps(attrs, indent=1))
    pad = np.zeros((chunk_rows, ns), np.float32)
    for i in range((n + chunk_rows - 1) // chunk_rows):
        r0, r1 = i * chunk_rows, min((i + 1) * chunk_rows, n)
        blk = data[r0:r1]
        blk = blk.detach().cpu().numpy() if hasattr(blk, "detach") else np.asarray(blk)
        pad[: r1 - r0] = blk
        pad[r1 - r0:] = 0.0
        with open(path / f"{i}.0", "wb") as fh:
            fh.write(pad.tobytes())
    return ZarrV2Array.open(path)


_ND_DTYPES = {"<f4": np.float32, "<f8": np.float64, "<i4": np.int32}


class ZarrV2Writer:
    """An uncompressed N-D Zarr v2 array written chunk by chunk (the output
    volumes: image (nx, ny, nt) f32 and fold (nx, ny, nt) i32, chunked
    (chunk_x, chunk_y, chunk_t) as the reference writes them,
    pstm/pipeline/executor.py:2066-2101).  Chunk files are named
    ``"i.j.k"`` and hold the full chunk shape in C order; an edge chunk is
    padded with the fill value, as Zarr pads it."""

    def __init__(self, path, shape, chunks, dtype, attrs: dict | None = None):
        self.path = Path(path)
        self.shape = tuple(int(s) for s in shape)
        self.chunks = tuple(int(c) for c in chunks)
        if len(self.chunks) != len(self.shape) or any(c < 1 for c in self.chunks):
            raise ValueError(f"chunks {chunks} do not fit shape {shape}")
        self.dtype = np.dtype(dtype)
        code = {np.dtype(v): k for k, v in _ND_DTYPES.items()}.get(self.dtype)
        if code is None:
            raise ValueError(f"unsupported dtype {self.dtype}")
        self.path.mkdir(parents=True, exist_ok=True)
        fill = 0.0 if self.dtype.kind == "f" else 0
        meta = {"zarr_format": 2, "shape": list(self.shape), "chunks": list(self.chunks), "dtype": code,
                "compressor": None, "fill_value": fill, "order": "C", "filters": None}
        (self.path / ".zarray").write_text(json.dumps(meta, indent=1))
        if attrs:
            (self.path / ".zattrs").write_text(json.dumps(attrs, indent=1))

    @property
    def grid(self) -> tuple[int, ...]:
        """Chunks along each axis."""
        return tuple((s + c - 1) // c for s, c in zip(self.shape, self.chunks))

    def bounds(self, idx) -> tuple[slice, ...]:
        """Array region of chunk ``idx``."""
        return tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, self.chunks, self.shape))

    def write_chunk(self, idx, block) -> int:
        """Write chunk ``idx`` from ``block`` (the region's shape, any layout;
        broadcast views are fine).  Returns the bytes written.  Releases the
        GIL while writing, so several threads write in parallel."""
        region = tuple(b.stop - b.start for b in self.bounds(idx))
        block = np.asarray(block)
        if block.shape != region:
            raise ValueError(f"chunk {idx}: block shape {block.shape} != region {region}")
        if region == self.chunks:
            buf = np.ascontiguousarray(block, dtype=self.dtype)
        else:
            buf = np.zeros(self.chunks, self.dtype)
            buf[tuple(slice(0, r) for r in region)] = block
        with open(self.path / ".".join(str(i) for i in idx), "wb") as fh:
            fh.write(memoryview(buf).cast("B"))
        return buf.nbytes


def read_zarr_v2(path) -> tuple[np.ndarray, dict]:
    """An uncompressed N-D Zarr v2 array read whole, with its attributes
    (tests and small arrays; a missing chunk is the fill value)."""
    path = Path(path)
    meta = json.loads((path / ".zarray").read_text())
    if meta.get("zarr_format") != 2 or meta.get("compressor") is not None or meta.get("filters"):
        raise NotImplementedError(f"{path}: only uncompressed Zarr v2 arrays are read here")
    if meta.get("order", "C") != "C" or meta.get("dtype") not in _ND_DTYPES:
        raise NotImplementedError(f"{path}: order {meta.get('order')}, dtype {meta.get('dtype')}")
    sep = meta.get("dimension_separator", ".")
    shape, chunks = tuple(meta["shape"]), tuple(meta["chunks"])
    dt = np.dtype(_ND_DTYPES[meta["dtype"]])
    out = np.full(shape, meta.get("fill_value") or 0, dt)
    for idx in np.ndindex(*((s + c - 1) // c for s, c in zip(shape, chunks))):
        f = path / sep.join(str(i) for i in idx)
        if not f.exists():
            continue
        blk = np.fromfile(f, dtype=dt).reshape(chunks)
        region = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, shape))
        out[region] = blk[tuple(slice(0, r.stop - r.start) for r in region)]
    attrs_path = path / ".zattrs"
    attrs = json.loads(attrs_path.read_text()) if attrs_path.exists() else {}
    return out, attrs


#: header columns, as the reference's ColumnMapping defaults (pstm/config/models.py:135-151)
HEADER_COLUMNS = ("trace_idx", "SOU_X", "SOU_Y", "REC_X", "REC_Y")


def write_headers(path, source_x, source_y, receiver_x, receiver_y, extra: dict | None = None) -> None:
    """Trace headers as a Parquet file with the reference's default column
    names (plus optional extra columns, e.g. an offset-bin number)."""
    import pyarrow as pa
    import pyarrow.parquet as pq

    n = len(source_x)
    cols = {"trace_idx": np.arange(n, dtype=np.int64), "SOU_X": np.asarray(source_x, np.float64),
            "SOU_Y": np.asarray(source_y, np.float64), "REC_X": np.asarray(receiver_x, np.float64),
            "REC_Y": np.asarray(receiver_y, np.float64)}
    for k, v in (extra or {}).items():
        cols[k] = np.asarray(v)
    pq.write_table(pa.table(cols), str(path))


def read_headers(path) -> dict:
    """Columns of a header Parquet file as NumPy arrays, in trace_idx order."""
    import pyarrow.parquet as pq

    t = pq.read_table(str(path))
    missing = [c for c in HEADER_COLUMNS if c not in t.column_names]
    if missing:
        raise ValueError(f"{path}: missing required columns {missing}")
    cols = {c: t.column(c).to_numpy() for c in t.column_names}
    order = np.argsort(cols["trace_idx"], kind="stable")
    if not np.array_equal(cols["trace_idx"][order], np.arange(len(order))):
        raise ValueError(f"{path}: trace_idx must be 0 .. n-1")
    return {k: v[order] for k, v in cols.items()}
