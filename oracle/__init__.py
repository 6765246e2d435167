"""CPU oracle for the HOOI-ALS hot path -- TEST INFRASTRUCTURE ONLY.

A plain numpy (FP64) restatement of the reference algorithm, each function
citing the reference file:line it follows.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import it, and only as the checker or the
timed CPU baseline; the product path (libdiana_hip.so and the C++ host layer)
never calls it.
"""
