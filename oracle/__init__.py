"""CPU FP64 oracle of the Binary-Supernova hot path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker.  The product never does.
"""
