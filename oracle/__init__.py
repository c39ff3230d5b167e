"""Test infrastructure: the FP64 NumPy oracle of the PSC-IMB step.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
