"""Blosc (format version 1 / 2 headers, c-blosc 1.x) chunk decoder.

The reference writes its trace arrays as Zarr v2 with the numcodecs Blosc
codec: zstd at clevel 3 with bit-shuffle (pstm/data/zarr_reader.py:416-436).
Neither numcodecs nor a blosc binding is installed here, so a chunk is
decoded from the published container format:

  * a 16-byte header: version, compressor version, flags, type size,
    uncompressed bytes, block size, compressed bytes (little-endian uint32);
    flags bit 0 byte-shuffle, bit 1 stored uncompressed (memcpy), bit 2
    bit-shuffle, bit 4 blocks not split into type-size streams, bits 5-7 the
    codec (0 BloscLZ, 1 LZ4, 2 Snappy, 3 zlib, 4 zstd);
  * for a compressed chunk, one int32 start offset per block, then per block
    one or `typesize` streams, each an int32 compressed length and the
    stream's bytes (a length equal to the stream's uncompressed size means
    the bytes are stored as they are);
  * each decompressed block is un-shuffled: byte-shuffle stores byte j of
    every element contiguously (the bytes that do not fill a whole element
    stay in place); bit-shuffle (the bitshuffle transform of Masui et al.
    2015) stores, for byte j and bit k of the elements, one bit per element
    packed eight elements to a byte, least significant bit first; c-blosc
    applies it to the whole block when the block holds a multiple of eight
    elements and leaves other blocks as they are (header version 2), or to
    the largest multiple of eight elements and leaves the rest (version 3).

zstd (RFC 8878), LZ4 block, Snappy and zlib streams are decompressed by the
installed pyarrow codecs and Python's zlib; BloscLZ has no decoder here.
"""

from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass

import numpy as np

CODECS = {0: "blosclz", 1: "lz4", 2: "snappy", 3: "zlib", 4: "zstd"}
_MAX_SPLITS = 16
_MIN_BUFFERSIZE = 128


@dataclass
class BloscHeader:
    version: int
    version_lz: int
    flags: int
    typesize: int
    nbytes: int
    blocksize: int
    cbytes: int

    @property
    def codec(self) -> str:
        return CODECS.get(self.flags >> 5, f"unknown({self.flags >> 5})")

    @property
    def byte_shuffle(self) -> bool:
        return bool(self.flags & 0x1)

    @property
    def memcpyed(self) -> bool:
        return bool(self.flags & 0x2)

    @property
    def bit_shuffle(self) -> bool:
        return bool(self.flags & 0x4)

    @property
    def dont_split(self) -> bool:
        return bool(self.flags & 0x10)

    @property
    def n_blocks(self) -> int:
        return (self.nbytes + self.blocksize - 1) // self.blocksize if self.blocksize else 0


def parse_header(buf) -> BloscHeader:
    b = memoryview(buf)
    if len(b) < 16:
        raise ValueError("blosc chunk shorter than its 16-byte header")
    version, version_lz, flags, typesize = b[0], b[1], b[2], b[3]
    nbytes, blocksize, cbytes = struct.unpack_from("<III", b, 4)
    h = BloscHeader(version, version_lz, flags, typesize, nbytes, blocksize, cbytes)
    if version not in (1, 2, 3):
        raise ValueError(f"blosc header version {version} not supported (1, 2 or 3)")
    if cbytes > len(b):
        raise ValueError(f"blosc chunk truncated: header says {cbytes} bytes, have {len(b)}")
    if typesize < 1 or (nbytes and blocksize < 1):
        raise ValueError("blosc header: bad type size or block size")
    return h


def _codec_decompress(codec: str, data, size: int) -> np.ndarray:
    if codec == "zlib":
        out = zlib.decompress(bytes(data))
    elif codec in ("zstd", "lz4", "snappy"):
        import pyarrow as pa

        name = "lz4_raw" if codec == "lz4" else codec
        out = pa.decompress(bytes(data), decompressed_size=size, codec=name, asbytes=True)
    else:
        raise NotImplementedError(f"blosc codec {codec!r} has no decoder here")
    if len(out) != size:
        raise ValueError(f"blosc {codec} stream decompressed to {len(out)} bytes, expected {size}")
    return np.frombuffer(out, dtype=np.uint8)


def bitunshuffle(src: np.ndarray, typesize: int, n_elem: int) -> np.ndarray:
    """Inverse bit-shuffle of n_elem elements (n_elem % 8 == 0): src holds,
    for each byte j < typesize and bit k < 8, n_elem / 8 bytes whose bit i
    (least significant first) is bit k of byte j of element 8 m + i."""
    planes = src[: typesize * n_elem].reshape(typesize, 8, n_elem // 8)
    bits = np.unpackbits(planes, axis=2, bitorder="little")            # (j, k, element)
    return np.packbits(bits.transpose(2, 0, 1), axis=2, bitorder="little").reshape(-1)  # (element, j)


def byteunshuffle(src: np.ndarray, typesize: int, nbytes: int) -> np.ndarray:
    n_elem = nbytes // typesize
    body = src[: n_elem * typesize].reshape(typesize, n_elem).T.reshape(-1)
    return np.concatenate([body, src[n_elem * typesize: nbytes]])


def _unshuffle_block(h: BloscHeader, blk: np.ndarray) -> np.ndarray:
    if h.bit_shuffle:
        n = len(blk) // h.typesize
        if h.version <= 2:
            if n % 8 != 0:
                return blk
            return bitunshuffle(blk, h.typesize, n)
        m = n - n % 8
        head = bitunshuffle(blk, h.typesize, m) if m else blk[:0]
        return np.concatenate([head, blk[m * h.typesize:]])
    if h.byte_shuffle and h.typesize > 1:
        return byteunshuffle(blk, h.typesize, len(blk))
    return blk


def decompress(buf) -> np.ndarray:
    """The uncompressed bytes of one blosc chunk (uint8 array)."""
    h = parse_header(buf)
    b = np.frombuffer(buf, dtype=np.uint8)
    if h.memcpyed:
        if 16 + h.nbytes > len(b):
            raise ValueError("blosc memcpyed chunk truncated")
        return b[16:16 + h.nbytes].copy()
    nb = h.n_blocks
    starts = struct.unpack_from(f"<{nb}i", buf, 16)
    out = np.empty(h.nbytes, dtype=np.uint8)
    for i, st in enumerate(starts):
        bsize = min(h.blocksize, h.nbytes - i * h.blocksize)
        split = (not h.dont_split and h.typesize <= _MAX_SPLITS and bsize // h.typesize >= _MIN_BUFFERSIZE)
        nsplit = h.typesize if split else 1
        ssize = bsize // nsplit
        parts = []
        pos = st
        for s in range(nsplit):
            if pos + 4 > len(b):
                raise ValueError(f"blosc block {i}: stream {s} header past the chunk end")
            (clen,) = struct.unpack_from("<i", buf, pos)
            pos += 4
            if clen < 0 or pos + clen > len(b):
                raise ValueError(f"blosc block {i}: bad stream length {clen}")
            data = b[pos:pos + clen]
            pos += clen
            parts.append(data.copy() if clen == ssize else _codec_decompress(h.codec, data, ssize))
        blk = parts[0] if nsplit == 1 else np.concatenate(parts)
        if len(blk) != bsize:
            raise ValueError(f"blosc block {i}: {len(blk)} bytes, expected {bsize}")
        out[i * h.blocksize: i * h.blocksize + bsize] = _unshuffle_block(h, blk)
    return out
