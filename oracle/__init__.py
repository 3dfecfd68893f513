"""Test infrastructure only: CPU restatements of the reference's 2-D IRLS unwrap.

Nothing in the shipped package imports this directory.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the checker.

Modules
-------
irls_fp64      numpy/scipy restatement of insardev/insardev/utils_unwrap2d.py:1082-1766
               in float64 (the accuracy oracle).
irls_torch32   torch-CPU float32 restatement of the same function, op for op (the
               "reference CPU path": cycle-count oracle and CPU timing baseline).
post           the reference's host epilogue (k-median, 4-connected labelling,
               per-component circular-mean alignment, size relabel, un-crop).
"""
