"""FP64 CPU oracle for the vegas+ hot path -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package.  The product package vegas_amd never does.
"""
