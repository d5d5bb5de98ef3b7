"""Test infrastructure: FP64 CPU restatement of the reference pair sums.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package.  The product package never imports it.
"""
