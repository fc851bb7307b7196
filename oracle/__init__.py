"""TEST INFRASTRUCTURE: CPU oracle of the reference PSTM migration kernel.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker (or the timed CPU baseline) — never as
part of the product path.
"""
