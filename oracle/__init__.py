"""CPU oracle of the FUGAL alignment path: an FP64 restatement of the
reference's CPU code (SebastianRueClausen/cuGAL, cugal/*.py), used as the
checker of the HIP kernels.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import it, and only to check or to time the
CPU baseline; the library (cugal_amd) never calls it.

How each function is pinned (tests/test_oracle.py) is stated in its
docstring; none is checked only by restating its own formula.
"""
