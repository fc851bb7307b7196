"""Callers of the kernel boundary: spatial trace selection and the tile
executor (pstm/data/spatial_index.py, pstm/pipeline/executor.py)."""
