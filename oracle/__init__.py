"""CPU oracle for the SSFM + FFT-Poisson hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, to check results.  The shipped library never
imports it.
"""
