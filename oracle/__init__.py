"""CPU oracle for the SSFM + FFT-Poisson hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The shipped library never
imports it.
"""
