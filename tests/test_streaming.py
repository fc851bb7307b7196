"""Streaming from disk (GPU): a common-offset-binned job written as an
uncompressed Zarr v2 array + Parquet headers, streamed to HBM through pinned
buffers and migrated bin by bin as the bins arrive, equals the in-memory
migration of the same traces (fold bit for bit, image to FP32 summation
order)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_streamed_bins_equal_in_memory(tmp_path):
    import torch

    from pm_amd.io.zarr_v2 import write_headers, write_zarr_v2
    from pm_amd.kernels import HipKernel, KernelConfig
    from pm_amd.pipeline.executor import TileExecutor
    from pm_amd.pipeline.streaming import migrate_streamed
    from pm_amd.synthetic import bins_job

    k = HipKernel()
    k.initialize(KernelConfig())
    job = bins_job(n_bins=12, total=9000, n_samples=801, n_out=32, nt=300, tile=16, n_planes=6,
                   n_diffractors=10, device="cuda")
    bins = np.repeat(np.arange(12), np.diff(job.meta["bin_start"]))
    write_zarr_v2(tmp_path / "traces.zarr", job.amplitudes, chunk_rows=1000,
                  attrs={"sample_rate_ms": job.dt_ms, "start_time_ms": job.t_start_ms})
    write_headers(tmp_path / "headers.parquet", job.source_x, job.source_y, job.receiver_x, job.receiver_y,
                  extra={"BIN": bins})
    res = migrate_streamed(k, tmp_path / "traces.zarr", tmp_path / "headers.parquet", job.x_axis, job.y_axis,
                           job.t_axis_ms, job.config, cube=job.cube, tile=job.tile, n_buffers=3, n_threads=2)
    print("streamed:", res.groups, "groups,", res.stats.chunks, "chunks,", res.stats.bytes_read, "bytes,",
          f"{res.wall_s:.2f} s", res.stats.errors)
    assert res.groups == 12 and res.stats.chunks == -(-job.n_traces // 1000) and not res.stats.errors
    assert torch.equal(res.executor.traces.amplitudes, job.amplitudes)
    ex = TileExecutor(k, job, spatial_sort=False)
    plan = ex.plan()
    ex.reset()
    plan.run()
    img, fold = ex.assemble()
    torch.cuda.synchronize()
    assert torch.equal(res.fold, fold)
    rel = float(torch.linalg.norm((res.image - img).double()) / torch.linalg.norm(img.double()))
    print("streamed vs in-memory image rel-L2", rel)
    assert rel < 1e-6
    plan.destroy()
    k.cleanup()
