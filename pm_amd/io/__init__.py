"""Trace input: Zarr v2 arrays (uncompressed) and Parquet headers, streamed to
HBM through pinned host buffers."""
