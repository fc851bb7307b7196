"""Zarr chunks to HBM through pinned host buffers.

The reference reads traces on the host per tile (ZarrTraceReader.get_traces,
pstm/data/zarr_reader.py:237-301) with an LRU trace cache
(pstm/data/trace_cache.py:20-140) and one prefetch thread that loads the next
tile while the current one computes (pstm/pipeline/executor.py:1203-1253).
Here the traces go to device memory once: reader threads fill a small pool of
pinned (page-locked) host buffers straight from the chunk files, each filled
buffer is copied to its rows of a device tensor on a dedicated copy stream,
and the buffer returns to the pool when its copy has finished.  Disk reads,
PCIe copies and GPU compute overlap; every chunk's copy completion is a
device event that compute streams wait on.
"""

from __future__ import annotations

import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from pm_amd.io.zarr_v2 import ZarrV2Array


@dataclass
class StreamStats:
    bytes_read: int = 0
    chunks: int = 0
    read_s: float = 0.0          # summed over reader threads
    wall_s: float = 0.0          # first read to last copy issued
    errors: list = field(default_factory=list)


class ChunkStreamer:
    """Streams the chunks of a Zarr v2 trace array into ``dest`` (a device
    float32 tensor of shape (n_traces, n_samples)), asynchronously."""

    def __init__(self, arr: ZarrV2Array, dest, n_buffers: int = 8, n_threads: int = 4):
        import torch

        if tuple(dest.shape) != arr.shape or dest.dtype != torch.float32 or not dest.is_cuda:
            raise ValueError(f"destination must be a device float32 tensor of shape {arr.shape}")
        if arr.dtype != np.float32:
            raise NotImplementedError("streaming reads float32 arrays")
        self.arr = arr
        self.dest = dest
        self.device = dest.device
        self.copy_stream = torch.cuda.Stream(device=self.device)  # default priority
        self.pool: queue.Queue = queue.Queue()
        for _ in range(n_buffers):
            self.pool.put(torch.empty((arr.chunk_rows, arr.n_samples), dtype=torch.float32, pin_memory=True))
        self.events = [None] * arr.n_chunks
        self._ready = [threading.Event() for _ in range(arr.n_chunks)]
        self._exec = ThreadPoolExecutor(max_workers=n_threads, thread_name_prefix="zarr-read")
        self._lock = threading.Lock()
        self.stats = StreamStats()
        self._t0 = None

    def _load(self, i: int) -> None:
        import torch

        buf = self.pool.get()
        try:
            t = time.perf_counter()
            rows = self.arr.read_chunk_into(i, buf.numpy())
            dt = time.perf_counter() - t
            r0, r1 = self.arr.chunk_range(i)
            ev = torch.cuda.Event()
            with torch.cuda.stream(self.copy_stream):
                self.dest[r0:r1].copy_(buf[:rows], non_blocking=True)
                ev.record(self.copy_stream)
            self.events[i] = ev
            with self._lock:
                self.stats.bytes_read += rows * self.arr.n_samples * 4
                self.stats.chunks += 1
                self.stats.read_s += dt
                self.stats.wall_s = time.perf_counter() - self._t0
            self._ready[i].set()
            ev.synchronize()        # the pinned buffer is free once its copy is done
        except Exception as exc:    # surfaced by wait_rows
            with self._lock:
                self.stats.errors.append(f"chunk {i}: {exc!r}")
            self._ready[i].set()
        finally:
            self.pool.put(buf)

    def start(self, chunks=None) -> "ChunkStreamer":
        """Queue the reads of the given chunk indices (all, in order, by
        default)."""
        if self._t0 is None:
            self._t0 = time.perf_counter()
        for i in (range(self.arr.n_chunks) if chunks is None else chunks):
            self._exec.submit(self._load, i)
        return self

    def wait_rows(self, r0: int, r1: int, stream=None) -> None:
        """Make ``stream`` (default: the current stream) wait until rows
        [r0, r1) are in device memory.  Blocks the host only until their copies
        have been issued."""
        import torch

        if r1 <= r0:
            return
        s = stream or torch.cuda.current_stream(self.device)
        c = self.arr.chunk_rows
        for i in range(r0 // c, (r1 - 1) // c + 1):
            self._ready[i].wait()
            if self.events[i] is None:
                raise IOError(f"streaming chunk {i} failed: {self.stats.errors}")
            s.wait_event(self.events[i])

    def close(self) -> None:
        self._exec.shutdown(wait=True)
        self.copy_stream.synchronize()
