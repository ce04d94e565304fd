"""CPU oracle of the per-grain level-set timestep -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker or the timed CPU baseline.  The product never imports it.
"""
